// Host-only half of the C ABI: models from triplets or MPS files, the
// warmup CSV reader, the point and FVA writers and the error text. Plain C++
// with no device code, built with the host compiler, so that the parsers of
// untrusted files also build and run under AddressSanitizer and
// UndefinedBehaviorSanitizer on a machine without a GPU (Makefile target
// `sanitize`, exercised by tests/test_sanitize.py).
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/achr.h"
#include "achr_model.h"

using achr_host::fail;

namespace achr_host {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// Closes a FILE* that is still open when a reader leaves early (an error
// return, or an allocation failure turned into ACHR_E_NOMEM by no_throw).
struct FileGuard {
  FILE*& f;
  ~FileGuard() {
    if (f) std::fclose(f);
  }
};

}  // namespace achr_host

extern "C" {

const char* achr_last_error(void) { return achr_host::g_err.c_str(); }
const char* achr_version(void) { return "achr-mi355x 0.1.0 (gfx950)"; }

void achr_free(void* p) { std::free(p); }

// ------------------------------------------------------------------ model

static int build_model(int n, int m, int nnz, const int* rows, const int* cols, const double* vals,
                       const double* lb, const double* ub, achr_model** out) {
  if (!out) return fail(ACHR_E_ARG, "out is NULL");
  *out = nullptr;
  if (n <= 0 || m < 0 || nnz < 0) return fail(ACHR_E_ARG, "bad sizes n=%d m=%d nnz=%d", n, m, nnz);
  if (nnz > 0 && (!rows || !cols || !vals)) return fail(ACHR_E_ARG, "NULL triplet arrays");
  if (!lb || !ub) return fail(ACHR_E_ARG, "NULL bounds");
  std::vector<std::map<int, double>> rowmap(static_cast<size_t>(m));
  for (int k = 0; k < nnz; ++k) {
    if (rows[k] < 0 || rows[k] >= m || cols[k] < 0 || cols[k] >= n)
      return fail(ACHR_E_ARG, "entry %d out of range (row %d col %d)", k, rows[k], cols[k]);
    rowmap[rows[k]][cols[k]] += vals[k];
  }
  auto* M = new (std::nothrow) achr_model();
  if (!M) return fail(ACHR_E_NOMEM, "model allocation");
  M->n = n;
  M->m = m;
  M->rowptr.assign(static_cast<size_t>(m) + 1, 0);
  for (int i = 0; i < m; ++i) {
    for (const auto& kv : rowmap[i]) {
      if (kv.second == 0.0) continue;
      M->colidx.push_back(kv.first);
      M->vals.push_back(kv.second);
    }
    M->rowptr[i + 1] = static_cast<int>(M->colidx.size());
  }
  M->lb.assign(lb, lb + n);
  M->ub.assign(ub, ub + n);
  M->n_file = n;
  *out = M;
  return ACHR_OK;
}

int achr_model_create(int n_rxns, int n_mets, int nnz, const int* rows, const int* cols, const double* vals,
                      const double* lb, const double* ub, achr_model** out) {
  return achr_host::no_throw([&] { return build_model(n_rxns, n_mets, nnz, rows, cols, vals, lb, ub, out); });
}

void achr_model_destroy(achr_model* model) { delete model; }

int achr_model_dims(const achr_model* M, int* n, int* m, int* nnz, int* d) {
  if (!M) return fail(ACHR_E_ARG, "model is NULL");
  if (n) *n = M->n;
  if (m) *m = M->m;
  if (nnz) *nnz = static_cast<int>(M->vals.size());
  if (d) *d = M->d;
  return ACHR_OK;
}

int achr_model_bounds(const achr_model* M, double* lb, double* ub) {
  if (!M) return fail(ACHR_E_ARG, "model is NULL");
  if (lb) std::copy(M->lb.begin(), M->lb.end(), lb);
  if (ub) std::copy(M->ub.begin(), M->ub.end(), ub);
  return ACHR_OK;
}

int achr_model_csr(const achr_model* M, int* rowptr, int* colidx, double* vals) {
  if (!M) return fail(ACHR_E_ARG, "model is NULL");
  if (rowptr) std::copy(M->rowptr.begin(), M->rowptr.end(), rowptr);
  if (colidx) std::copy(M->colidx.begin(), M->colidx.end(), colidx);
  if (vals) std::copy(M->vals.begin(), M->vals.end(), vals);
  return ACHR_OK;
}

// -------------------------------------------------------------- MPS reader

static bool next_tokens(const char* line, std::vector<std::string>& tok) {
  tok.clear();
  const char* p = line;
  while (*p) {
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
    if (!*p) break;
    const char* q = p;
    while (*q && *q != ' ' && *q != '\t' && *q != '\r' && *q != '\n') ++q;
    tok.emplace_back(p, q);
    p = q;
  }
  return !tok.empty();
}

static bool parse_num(const std::string& s, double* v) {
  errno = 0;
  char* end = nullptr;
  *v = std::strtod(s.c_str(), &end);
  return end && *end == '\0' && errno != ERANGE;
}

// What an MPS file holds, before either view of its rows is taken.
struct MpsFile {
  std::vector<std::string> row_names, col_names;
  std::vector<char> row_type;  // 'E', 'L', 'G', or 'N' for a free row after the objective
  std::vector<double> rhs, range;  // range: NaN where the RANGES section gives none
  std::vector<int> rr, cc;
  std::vector<double> vv, lb, ub;
};

static int parse_mps(const char* path, MpsFile& F) {
  if (!path) return fail(ACHR_E_ARG, "NULL argument");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(ACHR_E_IO, "cannot open MPS file %s", path);
  achr_host::FileGuard guard{f};
  enum Sec { NONE, ROWS, COLUMNS, RHS, RANGES, BOUNDS, DONE } sec = NONE;
  std::map<std::string, int> row_index;  // constraint rows only
  std::string obj_row;
  std::map<std::string, int> col_index;
  std::vector<std::string> tok;
  std::vector<char> buf(1 << 16);
  int lineno = 0;
  auto col_of = [&](const std::string& name) {
    auto it = col_index.find(name);
    if (it != col_index.end()) return it->second;
    const int j = static_cast<int>(col_index.size());
    col_index.emplace(name, j);
    F.col_names.push_back(name);
    F.lb.push_back(0.0);
    F.ub.push_back(DBL_MAX);
    return j;
  };
  int rc = ACHR_OK;
  while (std::fgets(buf.data(), static_cast<int>(buf.size()), f)) {
    ++lineno;
    if (buf[0] == '*') continue;
    const bool header = buf[0] != ' ' && buf[0] != '\t' && buf[0] != '\n' && buf[0] != '\r';
    if (!next_tokens(buf.data(), tok)) continue;
    if (header) {
      const std::string& h = tok[0];
      if (h == "NAME") sec = NONE;
      else if (h == "ROWS") sec = ROWS;
      else if (h == "COLUMNS") sec = COLUMNS;
      else if (h == "RHS") sec = RHS;
      else if (h == "RANGES") sec = RANGES;
      else if (h == "BOUNDS") sec = BOUNDS;
      else if (h == "ENDATA") { sec = DONE; break; }
      else if (h == "OBJSENSE" || h == "OBJSENS") sec = NONE;
      else { rc = fail(ACHR_E_FORMAT, "%s:%d: unknown section %s", path, lineno, h.c_str()); break; }
      continue;
    }
    if (sec == ROWS) {
      if (tok.size() < 2) { rc = fail(ACHR_E_FORMAT, "%s:%d: bad ROWS line", path, lineno); break; }
      const std::string& type = tok[0];
      if (type == "N" && obj_row.empty()) { obj_row = tok[1]; continue; }
      if (type != "N" && type != "E" && type != "L" && type != "G") {
        rc = fail(ACHR_E_FORMAT, "%s:%d: bad row type %s", path, lineno, type.c_str());
        break;
      }
      const int i = static_cast<int>(row_index.size());
      row_index.emplace(tok[1], i);
      F.row_names.push_back(tok[1]);
      F.row_type.push_back(type[0]);
      F.rhs.push_back(0.0);
      F.range.push_back(NAN);
    } else if (sec == COLUMNS) {
      if (tok.size() >= 3 && (tok[1] == "'MARKER'" || tok[2] == "'MARKER'")) continue;
      if (tok.size() != 3 && tok.size() != 5) {
        rc = fail(ACHR_E_FORMAT, "%s:%d: bad COLUMNS line", path, lineno);
        break;
      }
      const int j = col_of(tok[0]);
      for (size_t k = 1; k + 1 < tok.size(); k += 2) {
        double v;
        if (!parse_num(tok[k + 1], &v)) { rc = fail(ACHR_E_FORMAT, "%s:%d: bad number", path, lineno); break; }
        if (tok[k] == obj_row) continue;
        auto it = row_index.find(tok[k]);
        if (it == row_index.end()) {
          rc = fail(ACHR_E_FORMAT, "%s:%d: unknown row %s", path, lineno, tok[k].c_str());
          break;
        }
        F.rr.push_back(it->second);
        F.cc.push_back(j);
        F.vv.push_back(v);
      }
      if (rc != ACHR_OK) break;
    } else if (sec == RHS || sec == RANGES) {
      // [set name] row value [row value]: the set name is absent when the
      // line has an even number of fields.
      const size_t first = tok.size() % 2 == 0 ? 0 : 1;
      if (tok.size() < 2 || tok.size() > 5) {
        rc = fail(ACHR_E_FORMAT, "%s:%d: bad %s line", path, lineno, sec == RHS ? "RHS" : "RANGES");
        break;
      }
      for (size_t k = first; k + 1 < tok.size(); k += 2) {
        double v;
        if (!parse_num(tok[k + 1], &v)) { rc = fail(ACHR_E_FORMAT, "%s:%d: bad number", path, lineno); break; }
        if (tok[k] == obj_row) continue;  // minus the objective's constant term
        auto it = row_index.find(tok[k]);
        if (it == row_index.end()) {
          rc = fail(ACHR_E_FORMAT, "%s:%d: unknown row %s", path, lineno, tok[k].c_str());
          break;
        }
        (sec == RHS ? F.rhs : F.range)[it->second] = v;
      }
      if (rc != ACHR_OK) break;
    } else if (sec == BOUNDS) {
      if (tok.size() < 3) { rc = fail(ACHR_E_FORMAT, "%s:%d: bad BOUNDS line", path, lineno); break; }
      const std::string& type = tok[0];
      const bool novalue = type == "FR" || type == "MI" || type == "PL" || type == "BV";
      // Bound name may be omitted when the line has one field fewer.
      std::string cname;
      double v = 0.0;
      if (novalue) {
        cname = tok.size() >= 3 ? tok[2] : tok[1];
      } else {
        if (tok.size() < 4) { rc = fail(ACHR_E_FORMAT, "%s:%d: missing bound value", path, lineno); break; }
        cname = tok[2];
        if (!parse_num(tok[3], &v)) { rc = fail(ACHR_E_FORMAT, "%s:%d: bad number", path, lineno); break; }
      }
      auto it = col_index.find(cname);
      if (it == col_index.end()) {
        rc = fail(ACHR_E_FORMAT, "%s:%d: unknown column %s", path, lineno, cname.c_str());
        break;
      }
      const int j = it->second;
      double& lo = F.lb[j];
      double& hi = F.ub[j];
      if (type == "UP") {
        hi = v;
        if (v < 0.0 && lo == 0.0) lo = -DBL_MAX;  // MPS convention for a negative UP bound
      } else if (type == "LO") {
        lo = v;
      } else if (type == "FX") {
        lo = v;
        hi = v;
      } else if (type == "FR") {
        lo = -DBL_MAX;
        hi = DBL_MAX;
      } else if (type == "MI") {
        lo = -DBL_MAX;
      } else if (type == "PL") {
        hi = DBL_MAX;
      } else if (type == "BV") {
        lo = 0.0;
        hi = 1.0;
      } else if (type == "LI") {
        lo = v;
      } else if (type == "UI") {
        hi = v;
      } else {
        rc = fail(ACHR_E_FORMAT, "%s:%d: bad bound type %s", path, lineno, type.c_str());
        break;
      }
    }
  }
  std::fclose(f);
  f = nullptr;
  if (rc != ACHR_OK) return rc;
  if (sec != DONE) return fail(ACHR_E_FORMAT, "%s: missing ENDATA", path);
  if (F.col_names.empty()) return fail(ACHR_E_FORMAT, "%s: no columns", path);
  return ACHR_OK;
}

// The reference's view (ACHRCuda.cu:563-636): every row after the objective,
// whatever its type, is a row of S, and the sampled set is S x = 0; RHS and
// RANGES are read and ignored.
static int read_mps(const char* path, achr_model** out) {
  if (!path || !out) return fail(ACHR_E_ARG, "NULL argument");
  *out = nullptr;
  MpsFile F;
  int rc = parse_mps(path, F);
  if (rc != ACHR_OK) return rc;
  const int n = static_cast<int>(F.col_names.size());
  const int m = static_cast<int>(F.row_names.size());
  rc = build_model(n, m, static_cast<int>(F.vv.size()), F.rr.data(), F.cc.data(), F.vv.data(), F.lb.data(),
                   F.ub.data(), out);
  if (rc != ACHR_OK) return rc;
  (*out)->names = std::move(F.col_names);
  return ACHR_OK;
}

// The rows as written, the view GLPK gives the reference's warmup generator
// (createWarmupPtsGLPK.c:219): an E row is S_i x = b_i, an L row S_i x <= b_i,
// a G row S_i x >= b_i, RANGES give the other side (MPS convention: E with
// R > 0 is [b, b + R], with R < 0 [b + R, b]; L is [b - |R|, b]; G is
// [b, b + |R|]), free rows after the objective constrain nothing. The result
// is again a set {x' : S' x' = 0, lb' <= x' <= ub'} whose first n_file
// coordinates range over exactly the polytope of the file:
//   - a row with lo < hi gets a slack column s_i with S_i x - s_i = 0 and
//     bounds [lo, hi], an infinite side replaced by the row's activity range
//     over the column bounds (which the row can never leave);
//   - an equality row with b_i != 0 gets the coefficient -b_i on one column
//     fixed at 1, appended last.
static int read_mps_rows(const char* path, achr_model** out) {
  if (!path || !out) return fail(ACHR_E_ARG, "NULL argument");
  *out = nullptr;
  MpsFile F;
  int rc = parse_mps(path, F);
  if (rc != ACHR_OK) return rc;
  const int n_file = static_cast<int>(F.col_names.size());
  const int m_file = static_cast<int>(F.row_names.size());
  constexpr double kInf = 1e30;  // bounds at or beyond this are infinite
  std::vector<double> act_lo(m_file, 0.0), act_hi(m_file, 0.0);
  for (size_t k = 0; k < F.vv.size(); ++k) {
    const double a = F.vv[k], l = F.lb[F.cc[k]], u = F.ub[F.cc[k]];
    if (a == 0.0) continue;
    const double lo_term = a > 0 ? (l <= -kInf ? -INFINITY : a * l) : (u >= kInf ? -INFINITY : a * u);
    const double hi_term = a > 0 ? (u >= kInf ? INFINITY : a * u) : (l <= -kInf ? INFINITY : a * l);
    act_lo[F.rr[k]] += lo_term;
    act_hi[F.rr[k]] += hi_term;
  }
  std::vector<int> row_new(m_file, -1), aux_row;
  std::vector<double> aux_lb, aux_ub, rhs_coef(m_file, 0.0);
  std::vector<std::string> aux_name;
  int m = 0;
  bool need_one = false;
  for (int i = 0; i < m_file; ++i) {
    const char t = F.row_type[i];
    if (t == 'N') continue;  // a free row
    const double b = F.rhs[i], R = F.range[i];
    const bool ranged = !std::isnan(R);
    double lo = -INFINITY, hi = INFINITY;
    if (t == 'E') {
      lo = hi = b;
      if (ranged && R > 0) hi = b + R;
      if (ranged && R < 0) lo = b + R;
    } else if (t == 'L') {
      hi = b;
      if (ranged) lo = b - std::fabs(R);
    } else {
      lo = b;
      if (ranged) hi = b + std::fabs(R);
    }
    row_new[i] = m++;
    if (lo == hi) {
      if (lo < act_lo[i] || lo > act_hi[i])
        return fail(ACHR_E_FORMAT, "%s: row %s cannot be met within the column bounds ([%g, %g] excludes %g)", path,
                    F.row_names[i].c_str(), act_lo[i], act_hi[i], lo);
      if (b != 0.0) {
        rhs_coef[i] = -b;
        need_one = true;
      }
      continue;
    }
    lo = std::max(lo, act_lo[i]);
    hi = std::min(hi, act_hi[i]);
    if (!std::isfinite(lo) || !std::isfinite(hi))
      return fail(ACHR_E_FORMAT, "%s: row %s is unbounded %s within the column bounds; ACHR needs a bounded set",
                  path, F.row_names[i].c_str(), std::isfinite(lo) ? "above" : "below");
    if (lo > hi)
      return fail(ACHR_E_FORMAT, "%s: row %s cannot be met within the column bounds ([%g, %g])", path,
                  F.row_names[i].c_str(), lo, hi);
    aux_row.push_back(row_new[i]);
    aux_lb.push_back(lo);
    aux_ub.push_back(hi);
    aux_name.push_back("slack:" + F.row_names[i]);
  }
  const int n_slack = static_cast<int>(aux_row.size());
  const int n = n_file + n_slack + (need_one ? 1 : 0);
  std::vector<int> rr, cc;
  std::vector<double> vv;
  for (size_t k = 0; k < F.vv.size(); ++k) {
    if (row_new[F.rr[k]] < 0) continue;
    rr.push_back(row_new[F.rr[k]]);
    cc.push_back(F.cc[k]);
    vv.push_back(F.vv[k]);
  }
  for (int k = 0; k < n_slack; ++k) {
    rr.push_back(aux_row[k]);
    cc.push_back(n_file + k);
    vv.push_back(-1.0);
  }
  if (need_one) {
    for (int i = 0; i < m_file; ++i)
      if (rhs_coef[i] != 0.0) {
        rr.push_back(row_new[i]);
        cc.push_back(n - 1);
        vv.push_back(rhs_coef[i]);
      }
    aux_row.push_back(-1);
    aux_lb.push_back(1.0);
    aux_ub.push_back(1.0);
    aux_name.push_back("rhs:1");
  }
  std::vector<double> lb(F.lb), ub(F.ub);
  lb.insert(lb.end(), aux_lb.begin(), aux_lb.end());
  ub.insert(ub.end(), aux_ub.begin(), aux_ub.end());
  rc = build_model(n, m, static_cast<int>(vv.size()), rr.data(), cc.data(), vv.data(), lb.data(), ub.data(), out);
  if (rc != ACHR_OK) return rc;
  achr_model* M = *out;
  M->names = std::move(F.col_names);
  M->names.insert(M->names.end(), aux_name.begin(), aux_name.end());
  M->n_file = n_file;
  M->aux_row = std::move(aux_row);
  return ACHR_OK;
}

int achr_model_read_mps(const char* path, achr_model** out) {
  return achr_host::no_throw([&] { return read_mps(path, out); });
}

int achr_model_read_mps_rows(const char* path, achr_model** out) {
  return achr_host::no_throw([&] { return read_mps_rows(path, out); });
}

int achr_model_file_columns(const achr_model* M, int* n_file) {
  if (!M || !n_file) return fail(ACHR_E_ARG, "NULL argument");
  *n_file = M->n_file;
  return ACHR_OK;
}

// Slack k of row r takes the row's activity over the file's columns (CSR
// order, the slack's own entry skipped); the right-hand-side column is 1.
int achr_model_complete_points(const achr_model* M, const double* x_file, int P, double* x) {
  if (!M || (P > 0 && (!x_file || !x))) return fail(ACHR_E_ARG, "NULL argument");
  if (P < 0) return fail(ACHR_E_ARG, "P=%d", P);
  const int nf = M->n_file, n = M->n;
  for (int p = 0; p < P; ++p) {
    const double* xi = x_file + static_cast<size_t>(p) * nf;
    double* xo = x + static_cast<size_t>(p) * n;
    std::copy(xi, xi + nf, xo);
    for (int k = 0; k < n - nf; ++k) {
      const int r = M->aux_row[k];
      double s = 1.0;
      if (r >= 0) {
        s = 0.0;
        for (int q = M->rowptr[r]; q < M->rowptr[r + 1]; ++q)
          if (M->colidx[q] < nf) s += M->vals[q] * xi[M->colidx[q]];
      }
      xo[nf + k] = s;
    }
  }
  return ACHR_OK;
}

int achr_write_fva_csv(const char* path, const achr_model* M, const double* fva) {
  if (!path || !M || !fva) return fail(ACHR_E_ARG, "NULL argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(ACHR_E_IO, "cannot open %s for writing", path);
  std::fprintf(f, "reaction,min,max\n");
  for (int i = 0; i < M->n_file; ++i) {
    const std::string name = i < static_cast<int>(M->names.size()) ? M->names[i] : "?";
    std::fprintf(f, "%s,%f,%f\n", name.c_str(), fva[i], fva[static_cast<size_t>(M->n) + i]);
  }
  if (std::fclose(f) != 0) return fail(ACHR_E_IO, "write error on %s", path);
  return ACHR_OK;
}

// -------------------------------------------------------------- null space

int achr_model_set_nullspace(achr_model* M, const double* N, int d) {
  if (!M) return fail(ACHR_E_ARG, "model is NULL");
  if (d < 0 || d > M->n || (d > 0 && !N)) return fail(ACHR_E_ARG, "bad null-space dimension %d", d);
  return achr_host::no_throw([&] {
    M->N.assign(N, N + static_cast<size_t>(M->n) * d);
    M->d = d;
    return ACHR_OK;
  });
}

int achr_model_get_nullspace(const achr_model* M, double* N) {
  if (!M) return fail(ACHR_E_ARG, "model is NULL");
  if (M->d < 0) return fail(ACHR_E_STATE, "null space not computed");
  if (N) std::copy(M->N.begin(), M->N.end(), N);
  return ACHR_OK;
}
}  // extern "C"

namespace achr_host {

// The warmup file (ACHRCuda.cu:387-412, :676-689): one line per reaction,
// the points' values separated by commas. The file is read whole and split
// into lines (trailing '\r' / '\n' dropped, empty lines skipped, each line
// cut at its first NUL as a C string would be); the lines are parsed by up to
// 16 threads, each value by strtod, straight into the column-major array. An
// error names the first bad line in file order, as a sequential parse would.
int read_warmup_csv(const char* path, int n_expected, double** warmup, int* W_out, int* n_out) {
  if (!path || !warmup || !W_out) return fail(ACHR_E_ARG, "NULL argument");
  *warmup = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(ACHR_E_IO, "cannot open warmup CSV %s", path);
  FileGuard guard{f};
  std::string text;
  {
    long size = -1;
    if (std::fseek(f, 0, SEEK_END) == 0) size = std::ftell(f);
    std::rewind(f);
    if (size > 0) text.reserve(static_cast<size_t>(size) + 1);
    std::vector<char> chunk(1 << 20);
    size_t got = 0;
    while ((got = std::fread(chunk.data(), 1, chunk.size(), f)) > 0) text.append(chunk.data(), got);
  }
  std::fclose(f);
  f = nullptr;
  // non-empty lines as [begin, end) with text[end] set to '\0'
  std::vector<std::pair<size_t, size_t>> lines;
  for (size_t pos = 0; pos < text.size();) {
    size_t nl = text.find('\n', pos);
    const size_t stop = nl == std::string::npos ? text.size() : nl;
    size_t e = stop;
    if (const void* z = std::memchr(text.data() + pos, '\0', stop - pos))
      e = static_cast<size_t>(static_cast<const char*>(z) - text.data());
    while (e > pos && (text[e - 1] == '\r' || text[e - 1] == '\n')) --e;
    if (e > pos) lines.emplace_back(pos, e);
    pos = stop + 1;
  }
  text.push_back('\0');
  for (auto& l : lines) text[l.second] = '\0';
  const int n = static_cast<int>(lines.size());
  // one line's values: count, and store them at w[k * n + i] for k < W
  enum { kOk = 0, kBadNumber, kBadChar, kBadCount };
  auto parse = [&](int i, double* w, int W, int* count) -> int {
    const char* p = text.c_str() + lines[static_cast<size_t>(i)].first;
    int k = 0;
    while (true) {
      char* end = nullptr;
      const double v = std::strtod(p, &end);
      if (end == p) return kBadNumber;
      if (w && k < W) w[static_cast<size_t>(k) * n + i] = v;
      ++k;
      while (*end == ' ' || *end == '\t') ++end;
      if (*end == ',') { p = end + 1; continue; }
      if (*end == '\0') break;
      return kBadChar;
    }
    *count = k;
    return kOk;
  };
  auto report = [&](int i, int err, int count, int W) {
    if (err == kBadNumber) return fail(ACHR_E_FORMAT, "%s: line %d: bad number", path, i + 1);
    if (err == kBadChar) return fail(ACHR_E_FORMAT, "%s: line %d: unexpected character", path, i + 1);
    return fail(ACHR_E_FORMAT, "%s: line %d has %d values, expected %d", path, i + 1, count, W);
  };
  if (n == 0) return fail(ACHR_E_FORMAT, "%s: empty warmup file", path);
  int W = 0;
  if (const int err = parse(0, nullptr, 0, &W)) return report(0, err, 0, 0);  // never kBadCount
  auto* w = static_cast<double*>(std::malloc(sizeof(double) * static_cast<size_t>(n) * W));
  if (!w) return fail(ACHR_E_NOMEM, "warmup allocation");
  // lines [lo, hi) per thread; each keeps its first bad line
  const unsigned hw = std::thread::hardware_concurrency();
  const int T = n < 64 ? 1 : static_cast<int>(std::max(1u, std::min(16u, hw)));
  std::vector<int> bad_line(static_cast<size_t>(T), -1), bad_err(static_cast<size_t>(T), 0),
      bad_count(static_cast<size_t>(T), 0);
  auto work = [&](int t) {
    const int lo = static_cast<int>(static_cast<long long>(n) * t / T);
    const int hi = static_cast<int>(static_cast<long long>(n) * (t + 1) / T);
    for (int i = lo; i < hi; ++i) {
      int count = 0;
      int err = parse(i, w, W, &count);
      if (err == kOk && count != W) err = kBadCount;
      if (err) {
        bad_line[static_cast<size_t>(t)] = i;
        bad_err[static_cast<size_t>(t)] = err;
        bad_count[static_cast<size_t>(t)] = count;
        return;
      }
    }
  };
  std::vector<std::thread> pool;
  int started = 1;
  try {
    for (int t = 1; t < T; ++t) {
      pool.emplace_back(work, t);
      ++started;
    }
  } catch (const std::exception&) {
  }
  for (int t = started; t < T; ++t) work(t);
  work(0);
  for (auto& th : pool) th.join();
  for (int t = 0; t < T; ++t) {
    if (bad_line[static_cast<size_t>(t)] < 0) continue;
    std::free(w);
    const size_t u = static_cast<size_t>(t);
    return report(bad_line[u], bad_err[u], bad_count[u], W);
  }
  if (n_expected > 0 && n != n_expected) {
    std::free(w);
    return fail(ACHR_E_FORMAT, "%s: %d lines but the model has %d reactions", path, n, n_expected);
  }
  *warmup = w;
  *W_out = W;
  if (n_out) *n_out = n;
  return ACHR_OK;
}

}  // namespace achr_host

extern "C" {

int achr_read_warmup_csv(const char* path, int n_expected, double** warmup, int* W_out) {
  return achr_host::no_throw([&] { return achr_host::read_warmup_csv(path, n_expected, warmup, W_out, nullptr); });
}

// The reference's File<k> text (ACHRCuda.cu:744-756): one line per
// reaction, P "%f" values separated by commas. Rows are formatted by up to
// 16 threads, each into its own buffer (the same snprintf per value, so the
// bytes are those of a sequential fprintf), and written in order; rows go in
// batches so that the buffers hold about 64 MB at most. Where a thread
// cannot be started its rows are formatted by the calling thread.
int achr_write_points_csv(const char* path, const double* points, int n, int P) {
  if (!path || !points || n <= 0 || P <= 0) return fail(ACHR_E_ARG, "bad arguments");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(ACHR_E_IO, "cannot open %s for writing", path);
  const unsigned hw = std::thread::hardware_concurrency();
  const int T = static_cast<int>(std::max(1u, std::min(16u, hw)));
  const size_t per_row = static_cast<size_t>(P) * 16;  // typical "%f," width
  const int batch = static_cast<int>(
      std::max<size_t>(static_cast<size_t>(T), std::min<size_t>(static_cast<size_t>(n), (size_t{64} << 20) / per_row)));
  std::vector<std::string> buf(static_cast<size_t>(T));
  std::atomic<bool> nomem{false};
  auto format_rows = [&](int t, int r0, int r1) {
    try {
      std::string& out = buf[static_cast<size_t>(t)];
      out.clear();
      char tmp[400];  // "%f" of the largest double: 309 digits, sign, point and 6 decimals
      for (int i = r0; i < r1; ++i) {
        const double* row = points + static_cast<size_t>(i) * P;
        for (int j = 0; j < P; ++j) {
          const int len = std::snprintf(tmp, sizeof(tmp), j + 1 < P ? "%f," : "%f\n", row[j]);
          out.append(tmp, static_cast<size_t>(len));
        }
      }
    } catch (const std::bad_alloc&) {
      nomem = true;
    }
  };
  bool ok = true;
  for (int b0 = 0; b0 < n && ok && !nomem; b0 += batch) {
    const int b1 = std::min(n, b0 + batch);
    const int rows = b1 - b0, tn = std::min(T, rows);
    auto lo = [&](int t) { return b0 + static_cast<int>(static_cast<long long>(rows) * t / tn); };
    std::vector<std::thread> pool;
    int started = 1;  // partition 0 is this thread's
    try {
      for (int t = 1; t < tn; ++t) {
        pool.emplace_back(format_rows, t, lo(t), lo(t + 1));
        ++started;
      }
    } catch (const std::exception&) {
    }
    for (int t = started; t < tn; ++t) format_rows(t, lo(t), lo(t + 1));
    format_rows(0, lo(0), lo(1));
    for (auto& th : pool) th.join();
    for (int t = 0; t < tn && ok && !nomem; ++t)
      ok = std::fwrite(buf[static_cast<size_t>(t)].data(), 1, buf[static_cast<size_t>(t)].size(), f) ==
           buf[static_cast<size_t>(t)].size();
  }
  ok = (std::fclose(f) == 0) && ok;
  if (nomem) return fail(ACHR_E_NOMEM, "out of host memory formatting %s", path);
  if (!ok) return fail(ACHR_E_IO, "write error on %s", path);
  return ACHR_OK;
}

namespace {
constexpr char kPointsMagic[8] = {'A', 'C', 'H', 'R', 'P', 'T', '0', '1'};
}

int achr_write_points_bin(const char* path, const double* points, int n, int P) {
  if (!path || !points || n <= 0 || P <= 0) return fail(ACHR_E_ARG, "bad arguments");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(ACHR_E_IO, "cannot open %s for writing", path);
  const int32_t dims[2] = {n, P};
  const size_t cnt = static_cast<size_t>(n) * P;
  bool ok = std::fwrite(kPointsMagic, 1, 8, f) == 8 && std::fwrite(dims, sizeof(int32_t), 2, f) == 2 &&
            std::fwrite(points, sizeof(double), cnt, f) == cnt;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(ACHR_E_IO, "write error on %s", path);
  return ACHR_OK;
}

int achr_read_points_bin(const char* path, double** points, int* n, int* P) {
  if (!path || !points || !n || !P) return fail(ACHR_E_ARG, "NULL argument");
  *points = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(ACHR_E_IO, "cannot read %s: %s", path, std::strerror(errno));
  char magic[8];
  int32_t dims[2];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kPointsMagic, 8) != 0 ||
      std::fread(dims, sizeof(int32_t), 2, f) != 2 || dims[0] <= 0 || dims[1] <= 0) {
    std::fclose(f);
    return fail(ACHR_E_FORMAT, "%s is not a point file", path);
  }
  const size_t cnt = static_cast<size_t>(dims[0]) * dims[1];
  if (cnt > (static_cast<size_t>(1) << 40)) {
    std::fclose(f);
    return fail(ACHR_E_FORMAT, "%s: %d x %d points is not a plausible size", path, dims[0], dims[1]);
  }
  double* buf = static_cast<double*>(std::malloc(sizeof(double) * cnt));
  if (!buf) {
    std::fclose(f);
    return fail(ACHR_E_NOMEM, "out of host memory");
  }
  const bool ok = std::fread(buf, sizeof(double), cnt, f) == cnt;
  std::fclose(f);
  if (!ok) {
    std::free(buf);
    return fail(ACHR_E_FORMAT, "%s is truncated", path);
  }
  *points = buf;
  *n = dims[0];
  *P = dims[1];
  return ACHR_OK;
}

}  // extern "C"
