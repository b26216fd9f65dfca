// faspxx_dropin.cpp -- the FASP++-interface classes over the gfx950 C ABI.
#include "faspxx_dropin.hxx"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <stdexcept>

namespace {

int g_device = 0;
fx_ctx g_ctx = nullptr;

// Device failures inside void-returning methods abort with the library's
// message, as FASPXX_ABORT does for the reference's unrecoverable errors.
void check(int rc, const char* what) {
  if (rc != 0) {
    std::fprintf(stderr, "### ABORT: %s failed (%d): %s\n", what, rc, fx_last_error());
    std::abort();
  }
}

}  // namespace

namespace faspxx_amd {

void SetDevice(int device) { g_device = device; }

fx_ctx Context() {
  if (!g_ctx) check(fx_ctx_create(g_device, &g_ctx), "fx_ctx_create");
  return g_ctx;
}

}  // namespace faspxx_amd

using faspxx_amd::Context;

// ---------------------------------------------------------------------------
// VEC
// ---------------------------------------------------------------------------

void VEC::Resize(USI n) {
  if (dev && fx_vec_size(dev) == n) {
    size = n;
    return;
  }
  if (borrowed) {
    std::fprintf(stderr, "### ABORT: a vector handed to LOP::Apply cannot change size\n");
    std::abort();
  }
  if (dev) fx_vec_destroy(dev);
  dev = nullptr;
  check(fx_vec_create(Context(), n, FX_F64, &dev), "fx_vec_create");
  size = n;
  host.clear();
  host_valid = false;
  device_stale = false;
}

void VEC::SyncToHost() const {
  if (host_valid) return;
  host.resize(size);
  if (size) check(fx_vec_download(dev, host.data()), "fx_vec_download");
  host_valid = true;
}

void VEC::SyncToDevice() const {
  if (!dev) check(fx_vec_create(Context(), size, FX_F64, &dev), "fx_vec_create");
  if (device_stale) {
    if (size) check(fx_vec_upload(dev, host.data()), "fx_vec_upload");
    device_stale = false;
  }
}

fx_vec VEC::Device() const {
  SyncToDevice();
  return dev;
}

void VEC::DeviceWritten() {
  host_valid = false;
  device_stale = false;
}

fx_vec VEC::DeviceOut() {
  SyncToDevice();
  host_valid = false;
  return dev;
}

VEC::VEC(const USI& n, const DBL& value) : size(0) { SetValues(n, value); }

VEC::VEC(const std::vector<DBL>& src) : size(0) { SetValues(src); }

VEC::VEC(const USI& n, const DBL* src) : size(0) { SetValues(n, src); }

VEC::VEC(const VEC& src) : size(0) {
  Resize(src.size);
  check(fx_vec_copy(src.Device(), dev), "fx_vec_copy");
}

VEC::VEC(fx_vec view, bool b) : size(fx_vec_size(view)), borrowed(b), dev(view) {}

VEC::~VEC() {
  if (dev && !borrowed) fx_vec_destroy(dev);
}

VEC& VEC::operator=(const VEC& src) {
  if (this == &src) return *this;
  Resize(src.size);
  check(fx_vec_copy(src.Device(), dev), "fx_vec_copy");
  host_valid = false;
  return *this;
}

DBL& VEC::operator[](const USI& i) {
  SyncToHost();
  device_stale = true;
  return host[i];
}

const DBL& VEC::operator[](const USI& i) const {
  SyncToHost();
  return host[i];
}

VEC& VEC::operator+=(const VEC& v) {
  check(fx_add(Context(), v.Device(), DeviceOut()), "fx_add");
  return *this;
}

VEC& VEC::operator-=(const VEC& v) {
  check(fx_sub(Context(), v.Device(), DeviceOut()), "fx_sub");
  return *this;
}

void VEC::Reserve(const USI&) {}

void VEC::SetValues(const USI& n, const DBL& value) {
  Resize(n);
  check(fx_vec_set(dev, value), "fx_vec_set");
}

void VEC::SetValues(const std::vector<DBL>& src) {
  Resize((USI)src.size());
  if (size) check(fx_vec_upload(dev, src.data()), "fx_vec_upload");
}

void VEC::SetValues(const USI& n, const DBL* array) {
  if (array == nullptr || n == 0) {  // src/VEC.cxx:115-119
    Resize(0);
    return;
  }
  Resize(n);
  check(fx_vec_upload(dev, array), "fx_vec_upload");
}

DBL VEC::GetValue(const USI& position) const {
  SyncToHost();
  return host.at(position);
}

void VEC::GetValues(const USI& n, const USI* index, DBL* array) const {
  if (n == 0 || size == 0) return;  // src/VEC.cxx:136-139
  SyncToHost();
  for (USI j = 0; j < n; ++j) array[j] = host[index[j] % size];
}

void VEC::GetArray(DBL** array) {
  SyncToHost();
  device_stale = true;  // the caller may write through the pointer
  *array = host.data();
}

void VEC::GetArray(const DBL** array) const {
  SyncToHost();
  *array = host.data();
}

void VEC::Scale(const DBL& a) { check(fx_scale(Context(), a, DeviceOut()), "fx_scale"); }
void VEC::Reciprocal() { check(fx_recip(Context(), DeviceOut()), "fx_recip"); }
void VEC::PointwiseMult(const VEC& v) { check(fx_pmul(Context(), v.Device(), DeviceOut()), "fx_pmul"); }
void VEC::PointwiseDivide(const VEC& v) { check(fx_pdiv(Context(), v.Device(), DeviceOut()), "fx_pdiv"); }
void VEC::CopyTo(VEC& dst) const { dst = *this; }
void VEC::Shift(const DBL& a) { check(fx_shift(Context(), a, DeviceOut()), "fx_shift"); }
void VEC::Abs() { check(fx_abs(Context(), DeviceOut()), "fx_abs"); }
void VEC::AXPY(const DBL& a, const VEC& x) { check(fx_axpy(Context(), a, x.Device(), DeviceOut()), "fx_axpy"); }
void VEC::XPAY(const DBL& a, const VEC& x) { check(fx_xpay(Context(), a, x.Device(), DeviceOut()), "fx_xpay"); }
void VEC::AXPBY(const DBL& a, const DBL& b, const VEC& y) {
  check(fx_axpby(Context(), a, b, y.Device(), DeviceOut()), "fx_axpby");
}
void VEC::WAXPBY(const DBL& a, const VEC& x, const DBL& b, const VEC& y) {
  fx_vec xd = x.Device(), yd = y.Device();
  if (this != &x && this != &y) Resize(x.size);  // this->size = v1.size (src/VEC.cxx:332)
  check(fx_waxpby(Context(), a, xd, b, yd, DeviceOut()), "fx_waxpby");
}
DBL VEC::Max() const {
  double r = 0;
  check(fx_max(Context(), Device(), &r), "fx_max");
  return r;
}
DBL VEC::Min() const {
  double r = 0;
  check(fx_min(Context(), Device(), &r), "fx_min");
  return r;
}
DBL VEC::Norm2() const {
  double r = 0;
  check(fx_norm2(Context(), Device(), &r), "fx_norm2");
  return r;
}
DBL VEC::NormInf() const {
  double r = 0;
  check(fx_norminf(Context(), Device(), &r), "fx_norminf");
  return r;
}
DBL VEC::Dot(const VEC& v) const {
  double r = 0;
  check(fx_dot(Context(), Device(), v.Device(), &r), "fx_dot");
  return r;
}

// ---------------------------------------------------------------------------
// LOP (src/LOP.hxx:56-65: abort unless overridden)
// ---------------------------------------------------------------------------
void LOP::Apply(const VEC&, VEC&) const {
  std::fprintf(stderr, "### ABORT: Should be over-written!\n");
  std::abort();
}
void LOP::Residual(const VEC&, const VEC&, VEC&) const {
  std::fprintf(stderr, "### ABORT: Should be over-written!\n");
  std::abort();
}

// ---------------------------------------------------------------------------
// MAT
// ---------------------------------------------------------------------------
void MAT::Release() {
  if (csr) fx_csr_destroy(csr);
  csr = nullptr;
}

void MAT::Build(USI nr, USI mc, USI nz, const DBL* values, const USI* colInd, const USI* rowPtr,
                const USI* diagPtr) {
  Release();
  if (nr == 0 || mc == 0 || nz == 0) {  // MAT::Empty, src/MAT.cxx:26-29
    nrow = mcol = nnz = 0;
    has_values = false;
    return;
  }
  nrow = nr;
  mcol = mc;
  nnz = nz;
  has_values = values != nullptr;
  check(fx_csr_create(Context(), nr, mc, nz, rowPtr, colInd, values, diagPtr, &csr),
        "fx_csr_create");
}

MAT::MAT(const USI& nr, const USI& mc, const USI& nz, const std::vector<DBL>& values,
         const std::vector<USI>& colInd, const std::vector<USI>& rowPtr,
         const std::vector<USI>& diagPtr) {
  Build(nr, mc, nz, values.data(), colInd.data(), rowPtr.data(), diagPtr.data());
}

MAT::MAT(const USI& nr, const USI& mc, const USI& nz, const std::vector<DBL>& values,
         const std::vector<USI>& colInd, const std::vector<USI>& rowPtr) {
  Build(nr, mc, nz, values.data(), colInd.data(), rowPtr.data(), nullptr);
}

MAT::MAT(const USI& nr, const USI& mc, const USI& nz, const std::vector<USI>& colInd,
         const std::vector<USI>& rowPtr) {
  Build(nr, mc, nz, nullptr, colInd.data(), rowPtr.data(), nullptr);
}

MAT::MAT(const USI& nr, const USI& mc, const USI& nz, const std::vector<USI>& colInd,
         const std::vector<USI>& rowPtr, const std::vector<USI>& diagPtr) {
  Build(nr, mc, nz, nullptr, colInd.data(), rowPtr.data(), diagPtr.data());
}

static void diagonal_structure(USI n, std::vector<USI>& p) {
  p.resize(n + 1);
  for (USI j = 0; j <= n; ++j) p[j] = j;
}

MAT::MAT(const VEC& v) {
  const USI n = v.GetSize();
  std::vector<USI> p;
  diagonal_structure(n, p);
  std::vector<DBL> vals(n);
  if (n) check(fx_vec_download(v.Device(), vals.data()), "fx_vec_download");
  Build(n, n, n, vals.data(), p.data(), p.data(), p.data());
}

MAT::MAT(const std::vector<DBL>& v) {
  const USI n = (USI)v.size();
  std::vector<USI> p;
  diagonal_structure(n, p);
  Build(n, n, n, v.data(), p.data(), p.data(), p.data());
}

MAT::MAT(const MAT& m) : LOP() { *this = m; }

MAT::~MAT() { Release(); }

MAT& MAT::operator=(const MAT& m) {
  if (this == &m) return *this;
  if (!m.csr) {  // no entries: sizes only (an IdentityMatrix keeps its size)
    Release();
    nrow = m.nrow;
    mcol = m.mcol;
    nnz = 0;
    has_values = false;
    return *this;
  }
  std::vector<USI> rp(m.nrow + 1), ci(m.nnz), dp(std::min(m.nrow, m.mcol));
  std::vector<DBL> va(m.has_values ? m.nnz : 0);
  check(fx_csr_download(m.csr, rp.data(), ci.data(), m.has_values ? va.data() : nullptr,
                        dp.data()),
        "fx_csr_download");
  Build(m.nrow, m.mcol, m.nnz, m.has_values ? va.data() : nullptr, ci.data(), rp.data(),
        dp.data());
  return *this;
}

void MAT::SetValues(const USI& nr, const USI& mc, const USI& nz, const std::vector<DBL>& values,
                    const std::vector<USI>& colInd, const std::vector<USI>& rowPtr,
                    const std::vector<USI>& diagPtr) {
  Build(nr, mc, nz, values.data(), colInd.data(), rowPtr.data(), diagPtr.data());
}

void MAT::SetValues(const USI& nr, const USI& mc, const USI& nz, const std::vector<DBL>& values,
                    const std::vector<USI>& colInd, const std::vector<USI>& rowPtr) {
  Build(nr, mc, nz, values.data(), colInd.data(), rowPtr.data(), nullptr);
}

void MAT::GetDiag(VEC& v) const {
  v.Resize(nrow);
  if (nrow == 0) return;
  check(fx_csr_get_diag(Context(), csr, v.DeviceOut()), "fx_csr_get_diag");
}

void MAT::Scale(const DBL a) {
  if (!has_values || nnz == 0) return;  // structure only: src/MAT.cxx:353
  fx_vec vals = nullptr;
  check(fx_vec_wrap(Context(), nnz, FX_F64, fx_csr_values_dev(csr), &vals), "fx_vec_wrap");
  check(fx_scale(Context(), a, vals), "fx_scale");
  fx_vec_destroy(vals);
  check(fx_csr_set_values_dev(csr, fx_csr_values_dev(csr)), "fx_csr_set_values_dev");
}

void MAT::Apply(const VEC& v, VEC& w) const {
  if (nrow == 0) return;
  check(fx_spmv(Context(), csr, v.Device(), w.DeviceOut()), "MAT::Apply");
}

void MAT::Residual(const VEC& b, const VEC& x, VEC& r) const {
  if (nrow == 0) return;
  check(fx_residual(Context(), csr, b.Device(), x.Device(), r.DeviceOut()), "MAT::Residual");
}

void MAT::Adopt(fx_csr h, bool values) {
  Release();
  uint32_t nr = 0, mc = 0, nz = 0;
  check(fx_csr_info(h, &nr, &mc, &nz), "fx_csr_info");
  if (nz == 0) {  // MAT::Empty
    fx_csr_destroy(h);
    nrow = mcol = nnz = 0;
    has_values = false;
    return;
  }
  csr = h;
  nrow = nr;
  mcol = mc;
  nnz = nz;
  has_values = values;
}

void MAT::GetDiagInv(MAT& m) const {  // src/MAT.cxx:270-286: 1 / values[diagPtr[j]]
  VEC d;
  GetDiag(d);
  d.Reciprocal();
  m = MAT(d);
}

void MAT::GetLowerTri(MAT& lTri) const {
  if (!csr) return lTri.Empty();
  fx_csr h = nullptr;
  check(fx_csr_lower_tri(Context(), csr, &h), "fx_csr_lower_tri");
  lTri.Adopt(h, has_values);
}

void MAT::GetUpperTri(MAT& uTri) const {
  if (!csr) return uTri.Empty();
  fx_csr h = nullptr;
  check(fx_csr_upper_tri(Context(), csr, &h), "fx_csr_upper_tri");
  uTri.Adopt(h, has_values);
}

void MAT::Shift(const DBL a) {
  if (!csr || !has_values) return;
  check(fx_csr_shift_diag(Context(), csr, a), "fx_csr_shift_diag");
}

void MAT::Zero() {  // src/MAT.cxx:367-371
  if (!csr || !has_values) return;
  fx_vec vals = nullptr;
  check(fx_vec_wrap(Context(), nnz, FX_F64, fx_csr_values_dev(csr), &vals), "fx_vec_wrap");
  check(fx_vec_set(vals, 0.0), "fx_vec_set");
  fx_vec_destroy(vals);
  check(fx_csr_set_values_dev(csr, fx_csr_values_dev(csr)), "fx_csr_set_values_dev");
}

double* MAT::GetValues() const {
  double* v = new double[nnz ? nnz : 1];
  if (csr && has_values) check(fx_csr_download(csr, nullptr, nullptr, v, nullptr), "fx_csr_download");
  return v;
}

int* MAT::GetRowPtr() const {
  int* r = new int[(size_t)nrow + 1]();
  if (csr) check(fx_csr_download(csr, (uint32_t*)r, nullptr, nullptr, nullptr), "fx_csr_download");
  return r;
}

int* MAT::GetColInd() const {
  int* c = new int[nnz ? nnz : 1];
  if (csr) check(fx_csr_download(csr, nullptr, (uint32_t*)c, nullptr, nullptr), "fx_csr_download");
  return c;
}

void MAT::Add(const DBL a, const MAT& mat1, const DBL b, const MAT& mat2) {
  if (mat1.nnz == 0 || mat2.nnz == 0) return;
  fx_csr h = nullptr;
  check(fx_csr_add(Context(), a, mat1.csr, b, mat2.csr, &h), "fx_csr_add");
  Adopt(h, true);
}

void MAT::Mult(const MAT& matl, const MAT& matr) {
  if (!matl.csr || !matr.csr) return Empty();
  fx_csr h = nullptr;
  check(fx_csr_mult(Context(), matl.csr, matr.csr, &h), "fx_csr_mult");
  Adopt(h, matl.has_values && matr.has_values);
}

void MAT::MultLeft(const MAT& mat) {
  MAT tmp(*this);
  Mult(tmp, mat);
}

void MAT::MultRight(const MAT& mat) {
  MAT tmp(*this);
  Mult(mat, tmp);
}

void MAT::Empty() {
  Release();
  nrow = mcol = nnz = 0;
  has_values = false;
}

IdentityMatrix::IdentityMatrix(USI size) {
  nrow = size;
  mcol = size;
}

void MAT::TransInPlace() {
  if (!csr) return;
  fx_csr t = nullptr;
  check(fx_csr_transpose(Context(), csr, &t), "fx_csr_transpose");
  Release();
  csr = t;
  std::swap(nrow, mcol);
}

DBL MAT::GetValue(const USI& row, const USI& col) const {
  if (!csr || row >= nrow) return 0.0;
  std::vector<USI> rp(nrow + 1), ci(nnz);
  std::vector<DBL> va(nnz);
  check(fx_csr_download(csr, rp.data(), ci.data(), has_values ? va.data() : nullptr, nullptr),
        "fx_csr_download");
  for (USI k = rp[row]; k < rp[row + 1]; ++k)
    if (ci[k] == col) return has_values ? va[k] : 1.0;
  return 0.0;
}

// ---------------------------------------------------------------------------
// SOL, Identity, Jacobi
// ---------------------------------------------------------------------------
SOL::~SOL() {
  if (pc_cache) fx_pc_destroy(pc_cache);
}

FaspRetCode SOL::Setup(const LOP&) {
  std::fprintf(stderr, "### ABORT: Should be over-written!\n");
  std::abort();
}

FaspRetCode SOL::Solve(const VEC&, VEC&) {
  std::fprintf(stderr, "### ABORT: Should be over-written!\n");
  std::abort();
}

// SOL::SetSolTypeFromName, src/SOL.cxx:127-161
void SOL::SetSolTypeFromName(SOLParams& p) {
  for (char& c : p.algName) c = (char)std::tolower((unsigned char)c);
  static const struct {
    const char* name;
    SOLType type;
  } table[] = {{"cg", SOLVER_CG},         {"bicgstab", SOLVER_BICGSTAB}, {"minres", SOLVER_MINRES},
               {"gmres", SOLVER_GMRES},   {"fgmres", SOLVER_FGMRES},     {"vfgmres", SOLVER_VFGMRES},
               {"jacobi", SOLVER_JACOBI}, {"gs", SOLVER_GS},             {"sgs", SOLVER_SGS},
               {"sor", SOLVER_SOR},       {"ssor", SOLVER_SSOR},         {"mg", SOLVER_MG},
               {"fmg", SOLVER_FMG}};
  for (const auto& e : table) {
    if (p.algName == e.name) {
      p.type = e.type;
      return;
    }
  }
  p.type = SOLVER_CG;
  if (p.verbose > PRINT_NONE) std::cerr << "### WARNING: Unknown solver type. Using default solver!\n";
}

const char* SOL::GetSolType(SOLType type) const {
  switch (type) {
    case SOLVER_CG: return "CG";
    case SOLVER_BICGSTAB: return "BiCGStab";
    case SOLVER_MINRES: return "MinRes";
    case SOLVER_GMRES: return "GMRES";
    case SOLVER_FGMRES: return "FGMRES";
    case SOLVER_VFGMRES: return "VFGMRES";
    case SOLVER_JACOBI: return "JACOBI";
    case SOLVER_GS: return "GS";
    case SOLVER_SGS: return "SGS";
    case SOLVER_SOR: return "SOR";
    case SOLVER_SSOR: return "SSOR";
    case SOLVER_MG: return "MG";
    case SOLVER_FMG: return "FMG";
    default: std::fprintf(stderr, "### ABORT: Unknown solver type!\n"); std::abort();
  }
}

fx_pc SOL::DevicePC(const MAT&) { return nullptr; }

fx_params SOL::DeviceParams() const {
  fx_params p;
  fx_params_default(&p);
  p.maxIter = params.maxIter;
  p.minIter = params.minIter;
  p.savIter = params.savIter;
  p.restart = params.restart;
  p.relTol = params.relTol;
  p.absTol = params.absTol;
  p.verbose = params.verbose;
  return p;
}

void SOL::Record(const fx_result& r) {
  numIter = r.numIter;
  norm2 = r.norm2;
  normInf = r.normInf;
  // the closing report of SOL::PrintFinal, src/SOL.cxx:67-96
  if (params.minIter >= params.maxIter) return;
  if (params.verbose >= PRINT_MIN) {
    std::printf("---------------------------------------------\n");
    if (numIter >= params.maxIter) std::printf("### WARNING: maxIter = %u reached!\n", params.maxIter);
    std::printf("Number of iterations : %u\nNorm2 of residual    : %.5e\n"
                "NormInf of residual  : %.5e\n",
                numIter, norm2, normInf);
  } else if (params.verbose > PRINT_NONE) {
    std::printf("%s takes %u iterations to reach L2-norm of residual %.5e\n",
                GetSolType(params.type), numIter, norm2);
  }
}

FaspRetCode Identity::Solve(const VEC& b, VEC& x) {
  x = b;  // src/Iter.cxx:16-20
  return SUCCESS;
}

fx_pc Identity::DevicePC(const MAT&) {
  if (!pc_cache) check(fx_pc_create_identity(Context(), &pc_cache), "fx_pc_create_identity");
  return pc_cache;
}

Jacobi::~Jacobi() {
  if (own) fx_pc_destroy(own);
  if (solver) fx_solver_destroy(solver);
}

// Jacobi::Setup, src/Iter.cxx:29-53
FaspRetCode Jacobi::Setup(const MAT& A) {
  SetSolType(SOLVER_JACOBI);
  this->A = &A;
  mat = &A;
  if (own) fx_pc_destroy(own);
  own = nullptr;
  return SUCCESS;
}

fx_pc Jacobi::DevicePC(const MAT& A) {
  const MAT& M = mat ? *mat : A;
  const uint32_t iters = params.maxIter;
  if (!own || own_weight != weight || own_iters != iters) {
    if (own) fx_pc_destroy(own);
    own = nullptr;
    check(fx_pc_create_jacobi(Context(), M.Device(), weight, iters, &own), "fx_pc_create_jacobi");
    own_weight = weight;
    own_iters = iters;
  }
  return own;
}

// Jacobi::Solve, src/Iter.cxx:56-112.  With minIter == maxIter (the
// preconditioner form) it is a fixed number of sweeps and computes no norms.
FaspRetCode Jacobi::Solve(const VEC& b, VEC& x) {
  if (!mat) return ERROR_INPUT_PAR;
  fx_pc p = DevicePC(*mat);
  if (params.minIter == params.maxIter) {
    check(fx_pc_apply(Context(), p, b.Device(), x.DeviceOut()), "fx_pc_apply");
    numIter = params.maxIter;
    return SUCCESS;
  }
  if (!solver) check(fx_solver_create(Context(), FX_SOLVER_JACOBI, &solver), "fx_solver_create");
  const fx_params prm = DeviceParams();
  int rc = fx_solver_setup(solver, mat->Device(), p, &prm);
  fx_result res{};
  if (!rc) rc = fx_solver_solve(solver, b.Device(), x.DeviceOut(), &res);
  if (rc == FX_ERROR_DEVICE) check(rc, "Jacobi::Solve");
  Record(res);
  return (FaspRetCode)rc;
}

// ---------------------------------------------------------------------------
// Krylov solvers
// ---------------------------------------------------------------------------
namespace {

// y = A x for an operator that is not a MAT: its own Apply on views of the
// solver's device vectors; values it wrote through GetArray / operator[]
// go back to the device before the solver continues.
int lop_apply(void* user, fx_vec x, fx_vec y) {
  try {
    const LOP* L = static_cast<const LOP*>(user);
    const VEC xv(x, true);
    VEC yv(y, true);
    L->Apply(xv, yv);
    (void)yv.Device();
    return 0;
  } catch (...) {
    return ERROR_INPUT_PAR;
  }
}

const MAT& no_matrix() {
  static const MAT m;
  return m;
}

}  // namespace

KrylovSOL::~KrylovSOL() {
  if (op) fx_op_destroy(op);
  if (solver) fx_solver_destroy(solver);
}

FaspRetCode KrylovSOL::Setup(const LOP& A) {
  SetSolType(kind);
  this->A = &A;
  mat = dynamic_cast<const MAT*>(&A);
  lop = mat ? nullptr : &A;  // any other LOP: its Apply is called back
  setup_pending = true;  // the next Solve sets the device solver up afresh
  return SUCCESS;
}

// GMRES::Clean / FGMRES::Clean (src/GMRES.cxx:90-110, src/FGMRES.cxx:83-105):
// restart length back to the set-up value, FGMRES's Z zeroed.  CG and
// BiCGStab only zero work vectors every Solve overwrites.
void KrylovSOL::Clean() {
  if (solver) check(fx_solver_clean(solver), "fx_solver_clean");
}

FaspRetCode KrylovSOL::Solve(const VEC& b, VEC& x) {
  if (!mat && !lop) return ERROR_INPUT_PAR;
  fx_pc p = nullptr;
  if (pcd) {
    p = pcd->DevicePC(mat ? *mat : no_matrix());
    if (!p) return ERROR_SOLVER_PCD_TYPE;
  } else {
    if (!pc_cache) check(fx_pc_create_identity(Context(), &pc_cache), "fx_pc_create_identity");
    p = pc_cache;
  }
  if (!solver) check(fx_solver_create(Context(), (int)kind, &solver), "fx_solver_create");
  fx_params prm = DeviceParams();
  prm.maxRestart = maxRestart;
  prm.minRestart = minRestart;
  prm.leftPrecond = rightPrecond ? 0 : 1;
  int rc = 0;
  if (mat) {
    // the reference's Solve reads A, pcd and params as they are now; only a
    // Setup since the last Solve resets the GMRES / FGMRES restart length
    rc = setup_pending ? fx_solver_setup(solver, mat->Device(), p, &prm)
                       : fx_solver_update(solver, mat->Device(), p, &prm);
    setup_pending = false;
  } else {
    if (setup_pending || !op) {
      if (op) fx_op_destroy(op);
      op = nullptr;
      check(fx_op_create(Context(), lop->GetRowSize(), lop->GetColSize(), lop_apply,
                         const_cast<LOP*>(lop), &op),
            "fx_op_create");
      rc = fx_solver_setup_op(solver, op, p, &prm);
    } else {
      rc = fx_solver_update_op(solver, op, p, &prm);
    }
    setup_pending = false;
  }
  fx_result res{};
  if (!rc) rc = fx_solver_solve(solver, b.Device(), x.DeviceOut(), &res);
  if (rc == FX_ERROR_DEVICE) check(rc, "Solve");
  Record(res);
  return (FaspRetCode)rc;
}

// Krylov(), src/Krylov.cxx:16-53
FaspRetCode Krylov(LOP& A, VEC& b, VEC& x, SOL& pcd, SOLParams& params) {
  SOL::SetSolTypeFromName(params);
  std::unique_ptr<SOL> sol;
  switch (params.type) {
    case SOLVER_BICGSTAB: sol.reset(new BiCGStab()); break;
    case SOLVER_GMRES: sol.reset(new GMRES()); break;
    case SOLVER_FGMRES: sol.reset(new FGMRES()); break;
    default: sol.reset(new CG());
  }
  sol->SetOutput(params.verbose);
  sol->SetMaxIter(params.maxIter);
  sol->SetMinIter(params.minIter);
  sol->SetSavIter(params.savIter);
  sol->SetRestart(params.restart);
  sol->SetRelTol(params.relTol);
  sol->SetAbsTol(params.absTol);
  sol->Setup(A);
  sol->SetupPCD(pcd);
  return sol->Solve(b, x);
}
