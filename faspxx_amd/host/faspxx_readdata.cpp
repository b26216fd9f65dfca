// faspxx_readdata.cpp -- matrix and vector file ingestion with CSRx
// normalisation, for the drop-in classes: the reference's ReadVEC / ReadMTX /
// ReadCSR / ReadMat (src/ReadData.cxx:21-357) and SortCSRRow / MTXtoMAT /
// CSRtoMAT (src/MATUtil.cxx:252-476), with the same names, arguments and
// return codes.  Files are read on the host; ReadMat's result is a MAT, so
// the normalised matrix goes to HBM once.
//
// Formats (as the reference reads them):
//   .csr  n, then n+1 row pointers, nnz column indices, nnz values, one per
//         line; 1-based files (rowPtr[0] == 1) are shifted to 0-based
//   .mtx  '%' comment lines, "rows cols nnz", then nnz "i j value" triples,
//         1-based; the entries are taken as given (no symmetric expansion)
//   vec   n, then n values
// Normalisation: entries are bucketed by row in file order, each row's
// columns are sorted (stable insertion order for equal columns), and a zero
// diagonal is inserted in rows that have none.  Deviation: the reference's
// insertion sort walks an unsigned index below 0 when the first row needs
// reordering (src/MATUtil.cxx:452-458); here the sort is well defined for
// every row.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "ReadData.hxx"

namespace {

// Whole file as text; false if it cannot be opened.
bool slurp(const char* fileName, std::string& text) {
  std::ifstream in(fileName, std::ios::binary);
  if (!in.is_open()) return false;
  std::ostringstream ss;
  ss << in.rdbuf();
  text = ss.str();
  return true;
}

// Whitespace-separated token reader over a text buffer.
struct Tokens {
  const char* p;
  const char* end;
  explicit Tokens(const std::string& s) : p(s.data()), end(s.data() + s.size()) {}
  bool next(std::string& tok) {
    while (p < end && std::isspace((unsigned char)*p)) ++p;
    if (p >= end) return false;
    const char* b = p;
    while (p < end && !std::isspace((unsigned char)*p)) ++p;
    tok.assign(b, p);
    return true;
  }
  bool next_long(long long& v) {
    std::string t;
    if (!next(t)) return false;
    char* e = nullptr;
    errno = 0;
    v = std::strtoll(t.c_str(), &e, 10);
    return e && *e == '\0' && errno == 0;
  }
  bool next_double(double& v) {
    std::string t;
    if (!next(t)) return false;
    char* e = nullptr;
    v = std::strtod(t.c_str(), &e);
    return e && *e == '\0';
  }
  // skip lines starting with '%' at the current position
  void skip_comments() {
    for (;;) {
      while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
      if (p < end && *p == '%') {
        while (p < end && *p != '\n') ++p;
      } else {
        return;
      }
    }
  }
};

}  // namespace

FaspRetCode ReadVEC(const char* fileName, VEC& dst) {
  std::string text;
  if (!slurp(fileName, text)) return FaspRetCode::ERROR_OPEN_FILE;
  Tokens t(text);
  long long len = 0;
  if (!t.next_long(len) || len < 0) return FaspRetCode::ERROR_INPUT_FILE;
  std::vector<DBL> v;
  v.reserve((size_t)len);
  double x;
  while ((long long)v.size() < len && t.next_double(x)) v.push_back(x);
  v.resize((size_t)len, 0.0);
  dst.SetValues(v);
  return (long long)v.size() == len && t.p <= t.end ? FaspRetCode::SUCCESS
                                                      : FaspRetCode::ERROR_INPUT_FILE;
}

FaspRetCode ReadMTX(const char* fileName, USI& row, USI& col, USI& nnz, std::vector<USI>& rowInd,
                    std::vector<USI>& colInd, std::vector<DBL>& values) {
  std::string text;
  if (!slurp(fileName, text)) return FaspRetCode::ERROR_OPEN_FILE;
  Tokens t(text);
  t.skip_comments();
  long long r = 0, c = 0, z = 0;
  if (!t.next_long(r) || !t.next_long(c) || !t.next_long(z) || r <= 0 || c <= 0 || z < 0)
    return FaspRetCode::ERROR_INPUT_FILE;
  row = (USI)r;
  col = (USI)c;
  nnz = (USI)z;
  rowInd.assign(nnz, 0);
  colInd.assign(nnz, 0);
  values.assign(nnz, 0.0);
  for (USI k = 0; k < nnz; ++k) {
    long long i = 0, j = 0;
    double v = 0.0;
    if (!t.next_long(i) || !t.next_long(j) || !t.next_double(v)) return FaspRetCode::ERROR_INPUT_FILE;
    if (i < 1 || j < 1 || i > r || j > c) return FaspRetCode::ERROR_INPUT_FILE;
    rowInd[k] = (USI)(i - 1);
    colInd[k] = (USI)(j - 1);
    values[k] = v;
  }
  return FaspRetCode::SUCCESS;
}

FaspRetCode ReadCSR(const char* fileName, USI& row, USI& col, USI& nnz, std::vector<USI>& rowPtr,
                    std::vector<USI>& colInd, std::vector<DBL>& values) {
  std::string text;
  if (!slurp(fileName, text)) return FaspRetCode::ERROR_OPEN_FILE;
  Tokens t(text);
  long long n = 0;
  if (!t.next_long(n)) return FaspRetCode::ERROR_INPUT_FILE;
  if (n <= 0) return FaspRetCode::ERROR_INPUT_PAR;  // src/ReadData.cxx:274-278
  row = col = (USI)n;
  rowPtr.assign((size_t)n + 1, 0);
  for (long long i = 0; i <= n; ++i) {
    long long v = 0;
    if (!t.next_long(v) || v < 0) return FaspRetCode::ERROR_INPUT_FILE;
    rowPtr[(size_t)i] = (USI)v;
  }
  if (rowPtr[(size_t)n] < rowPtr[0]) return FaspRetCode::ERROR_INPUT_FILE;
  nnz = rowPtr[(size_t)n] - rowPtr[0];
  colInd.assign(nnz, 0);
  values.assign(nnz, 0.0);
  for (USI k = 0; k < nnz; ++k) {
    long long v = 0;
    if (!t.next_long(v) || v < 0) return FaspRetCode::ERROR_INPUT_FILE;
    colInd[k] = (USI)v;
  }
  USI got = 0;
  for (; got < nnz; ++got) {
    double v = 0.0;
    if (!t.next_double(v)) break;
    values[got] = v;
  }
  if (rowPtr[0] == 1) {  // 1-based file
    for (auto& p : rowPtr) --p;
    for (auto& c : colInd) --c;
  }
  return got == nnz ? FaspRetCode::SUCCESS : FaspRetCode::ERROR_INPUT_FILE;
}

// Structure checks of CheckCSR (src/MATUtil.cxx:110-175) on the arrays given.
static FaspRetCode check_csr(USI row, USI col, USI nnz, const std::vector<DBL>& values,
                             const std::vector<USI>& colInd, const std::vector<USI>& rowPtr) {
  if (row == 0 || col == 0 || nnz == 0) return FaspRetCode::SUCCESS;
  if (rowPtr.size() != (size_t)row + 1 || colInd.size() != nnz || values.size() != nnz ||
      rowPtr[row] - rowPtr[0] != nnz)
    return FaspRetCode::ERROR_MAT_DATA;
  for (USI j = 0; j < row; ++j)
    if (rowPtr[j] > rowPtr[j + 1]) return FaspRetCode::ERROR_MAT_DATA;
  for (USI k = 0; k < nnz; ++k)
    if (colInd[k] >= col) return FaspRetCode::ERROR_MAT_DATA;
  return FaspRetCode::SUCCESS;
}

FaspRetCode SortCSRRow(const USI& row, const USI& col, const USI& nnz,
                       const std::vector<USI>& rowPtr, std::vector<USI>& colInd,
                       std::vector<DBL>& values) {
  std::vector<std::pair<USI, DBL>> e;
  for (USI j = 0; j < row; ++j) {
    const USI b = rowPtr[j], en = rowPtr[j + 1];
    if (en <= b + 1) continue;
    e.clear();
    for (USI k = b; k < en; ++k) e.emplace_back(colInd[k], values[k]);
    std::stable_sort(e.begin(), e.end(),
                     [](const std::pair<USI, DBL>& a, const std::pair<USI, DBL>& c) {
                       return a.first < c.first;
                     });
    for (USI k = b; k < en; ++k) {
      colInd[k] = e[k - b].first;
      values[k] = e[k - b].second;
    }
  }
  return check_csr(row, col, nnz, values, colInd, rowPtr);
}

// Rows without a diagonal entry get a zero one, in column order
// (src/MATUtil.cxx:303-437): the CSRx arrays CSRtoMAT hands to MAT.
static FaspRetCode insert_zero_diag(USI row, USI col, USI nnz, const std::vector<DBL>& values,
                                    const std::vector<USI>& colInd,
                                    const std::vector<USI>& rowPtr, std::vector<USI>& rp,
                                    std::vector<USI>& ci, std::vector<DBL>& va) {
  const FaspRetCode rc = check_csr(row, col, nnz, values, colInd, rowPtr);
  if (rc != FaspRetCode::SUCCESS) return rc;
  rp.assign(1, 0);
  ci.clear();
  va.clear();
  ci.reserve(nnz + row);
  va.reserve(nnz + row);
  for (USI i = 0; i < row; ++i) {
    const USI b = rowPtr[i], en = rowPtr[i + 1];
    bool has = false;
    for (USI k = b; k < en; ++k) has = has || colInd[k] == i;
    bool put = has || i >= col;
    for (USI k = b; k < en; ++k) {
      if (!put && colInd[k] > i) {
        ci.push_back(i);
        va.push_back(0.0);
        put = true;
      }
      ci.push_back(colInd[k]);
      va.push_back(values[k]);
    }
    if (!put) {
      ci.push_back(i);
      va.push_back(0.0);
    }
    rp.push_back((USI)ci.size());
  }
  return FaspRetCode::SUCCESS;
}

// Bucket COO entries by row in file order (src/MATUtil.cxx:62-108).
static FaspRetCode coo_to_csr(USI row, USI col, USI nnz, const std::vector<USI>& rowInd,
                              const std::vector<USI>& colInd, const std::vector<DBL>& values,
                              std::vector<USI>& rp, std::vector<USI>& ci, std::vector<DBL>& va) {
  rp.assign(row + 1, 0);
  ci.assign(nnz, 0);
  va.assign(nnz, 0.0);
  for (USI k = 0; k < nnz; ++k) {
    if (rowInd[k] >= row || colInd[k] >= col) return FaspRetCode::ERROR_MAT_DATA;
    ++rp[rowInd[k] + 1];
  }
  for (USI i = 0; i < row; ++i) rp[i + 1] += rp[i];
  std::vector<USI> cur(rp.begin(), rp.end() - 1);
  for (USI k = 0; k < nnz; ++k) {
    const USI p = cur[rowInd[k]]++;
    ci[p] = colInd[k];
    va[p] = values[k];
  }
  return FaspRetCode::SUCCESS;
}

FaspRetCode CSRtoMAT(const USI& row, const USI& col, const USI& nnz,
                     const std::vector<DBL>& values, const std::vector<USI>& colInd,
                     const std::vector<USI>& rowPtr, MAT& mat) {
  std::vector<USI> rp, ci;
  std::vector<DBL> va;
  const FaspRetCode rc = insert_zero_diag(row, col, nnz, values, colInd, rowPtr, rp, ci, va);
  if (rc != FaspRetCode::SUCCESS) return rc;
  mat.SetValues(row, col, (USI)ci.size(), va, ci, rp);
  return FaspRetCode::SUCCESS;
}

FaspRetCode MTXtoMAT(const USI& row, const USI& col, const USI& nnz,
                     const std::vector<USI>& rowInd, const std::vector<USI>& colInd,
                     const std::vector<DBL>& values, MAT& mat) {
  std::vector<USI> rp, ci;
  std::vector<DBL> va;
  FaspRetCode rc = coo_to_csr(row, col, nnz, rowInd, colInd, values, rp, ci, va);
  if (rc != FaspRetCode::SUCCESS) return rc;
  rc = SortCSRRow(row, col, nnz, rp, ci, va);
  if (rc != FaspRetCode::SUCCESS) return rc;
  return CSRtoMAT(row, col, nnz, va, ci, rp, mat);
}

FaspRetCode ReadCSRx(const char* fileName, USI& row, USI& col, USI& nnz,
                     std::vector<USI>& rowPtr, std::vector<USI>& colInd,
                     std::vector<DBL>& values) {
  const size_t len = std::strlen(fileName);
  if (len <= 4) return FaspRetCode::ERROR_INPUT_FILE;
  std::string ext(fileName + len - 3);
  for (auto& ch : ext) ch = (char)std::tolower((unsigned char)ch);
  std::vector<USI> a, ci, rp;
  std::vector<DBL> va;
  FaspRetCode rc;
  if (ext == "csr") {
    rc = ReadCSR(fileName, row, col, nnz, rp, ci, va);
    if (rc != FaspRetCode::SUCCESS) return rc;
  } else if (ext == "mtx") {
    rc = ReadMTX(fileName, row, col, nnz, a, ci, va);
    if (rc != FaspRetCode::SUCCESS) return rc;
    std::vector<USI> c2;
    std::vector<DBL> v2;
    rc = coo_to_csr(row, col, nnz, a, ci, va, rp, c2, v2);
    if (rc != FaspRetCode::SUCCESS) return rc;
    ci.swap(c2);
    va.swap(v2);
  } else {
    return FaspRetCode::ERROR_INPUT_FILE;
  }
  rc = SortCSRRow(row, col, nnz, rp, ci, va);
  if (rc != FaspRetCode::SUCCESS) return rc;
  rc = insert_zero_diag(row, col, nnz, va, ci, rp, rowPtr, colInd, values);
  if (rc != FaspRetCode::SUCCESS) return rc;
  nnz = (USI)colInd.size();
  return FaspRetCode::SUCCESS;
}

FaspRetCode ReadMat(const char* fileName, MAT& dst) {
  USI row = 0, col = 0, nnz = 0;
  std::vector<USI> rp, ci;
  std::vector<DBL> va;
  const FaspRetCode rc = ReadCSRx(fileName, row, col, nnz, rp, ci, va);
  if (rc != FaspRetCode::SUCCESS) return rc;
  dst.SetValues(row, col, nnz, va, ci, rp);
  return FaspRetCode::SUCCESS;
}
