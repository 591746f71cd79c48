// gpufem -- C++ host classes with the reference's call surface, built on the
// C ABI of libgpufem_amd.so (include/gpufem_amd.h).
//
// Reference classes and functions mirrored here (names, argument meaning and
// conventions kept; errors throw gpufem::Error where the reference called
// exit()):
//   HostVec / DeviceVec / VecType        include/linalg/vec.h:146-395
//   BsrData                              include/linalg/bsr_utils.h:43-404
//   BsrMat                               include/linalg/bsr_mat.h:10-356
//   ILUk_Preconditioner                  include/linalg/precond.h:5-55
//   bsr_pre_solve / bsr_post_solve       include/linalg/solvers/solve_utils.h
//   ElementAssembler (+ createFromBDF)   include/assembler.h:14-321
//   TACSMeshLoader                       include/mesh/TACSMeshLoader.h:138-227
//   shell element type tags              include/shell/{quadrature,director,
//                                        basis,data,physics,elem_group}.h
//   TacsComputeRCMOrder, computeILUk,    include/linalg/reordering_utils.h:272,
//   getBandWidth, get_elem_ind_map       include/linalg/bsr_utils.h:655-870
//   createPlateAssembler, getPlateLoads  examples/plate_demo/_plate_utils.h
//   solvers::gmres (the flow of examples/iterative/gmres.cu, restarted and
//   right preconditioned) and solvers::direct_LU_solve
//   (include/linalg/solvers/cusparse.h:16 direct_LU_solve_old)
//
// Ownership differs from the reference on purpose: vectors and BSR data are
// shallow handles (copies alias, as in the reference) but reference-counted,
// so memory is released when the last handle goes.  All device work runs on
// one library context (device 0 unless gpufem::set_device is called first).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../gpufem_amd.h"

namespace gpufem {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline int& device_index() {
  static int d = 0;
  return d;
}
inline void set_device(int d) { device_index() = d; }

// process-wide library context (one stream)
inline gf_context context() {
  static std::shared_ptr<gf_context_s> ctx = [] {
    gf_context c = nullptr;
    int st = gf_context_create(device_index(), &c);
    if (st != GF_OK) throw Error(st, "gf_context_create failed");
    return std::shared_ptr<gf_context_s>(c, gf_context_destroy);
  }();
  return ctx.get();
}

inline void check(int st, const char* what) {
  if (st != GF_OK)
    throw Error(st, std::string(what) + ": " + gf_context_last_error(context()));
}
inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(GF_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
inline void synchronize() { check(gf_context_synchronize(context()), "sync"); }

template <typename T>
class DeviceVec;

// ------------------------------------------------------------------ vectors
template <typename T>
class HostVec {
 public:
  using type = T;
  HostVec() = default;
  explicit HostVec(int N) : N_(N), buf_(std::make_shared<std::vector<T>>(N)) {
    data_ = buf_->data();
  }
  HostVec(int N, T* data) : N_(N), data_(data) {}  // wraps caller memory
  HostVec(int N, T value) : HostVec(N) {
    for (int i = 0; i < N; i++) data_[i] = value;
  }
  T* getPtr() { return data_; }
  const T* getPtr() const { return data_; }
  int getSize() const { return N_; }
  T& operator[](int i) { return data_[i]; }
  const T& operator[](int i) const { return data_[i]; }
  void zeroValues() { std::memset(data_, 0, sizeof(T) * N_); }
  void scale(T s) {
    for (int i = 0; i < N_; i++) data_[i] *= s;
  }
  void randomize(T maxVal = T(1), unsigned seed = 0) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (int i = 0; i < N_; i++) data_[i] = T(maxVal * u(g));
  }
  HostVec copyVec() const {
    HostVec v(N_);
    std::memcpy(v.getPtr(), data_, sizeof(T) * N_);
    return v;
  }
  // temp[perm[node] * block + k] = data[node * block + k]
  void permuteData(int block_dim, const int* perm) {
    std::vector<T> tmp(N_);
    const int nn = N_ / block_dim;
    for (int i = 0; i < nn; i++)
      for (int k = 0; k < block_dim; k++)
        tmp[perm[i] * block_dim + k] = data_[i * block_dim + k];
    std::memcpy(data_, tmp.data(), sizeof(T) * N_);
  }
  HostVec createPermuteVec(int block_dim, const int* perm) const {
    HostVec v = copyVec();
    v.permuteData(block_dim, perm);
    return v;
  }
  void apply_bcs(const HostVec<int>& bcs) {
    for (int i = 0; i < bcs.getSize(); i++) data_[bcs[i]] = T(0);
  }
  DeviceVec<T> createDeviceVec(bool memset = true) const;
  HostVec createHostVec() const { return *this; }

 private:
  int N_ = 0;
  T* data_ = nullptr;
  std::shared_ptr<std::vector<T>> buf_;
};

template <typename T>
class DeviceVec {
 public:
  using type = T;
  DeviceVec() = default;
  explicit DeviceVec(int N, bool memset = true) : N_(N) {
    T* p = nullptr;
    hip_check(hipMalloc(&p, sizeof(T) * (N ? N : 1)), "hipMalloc");
    buf_ = std::shared_ptr<T>(p, [](T* q) { hipFree(q); });
    data_ = p;
    if (memset) zeroValues();
  }
  DeviceVec(int N, T* data) : N_(N), data_(data) {}  // wraps device memory
  T* getPtr() { return data_; }
  const T* getPtr() const { return data_; }
  int getSize() const { return N_; }
  void zeroValues() {
    hip_check(hipMemsetAsync(data_, 0, sizeof(T) * N_,
                             (hipStream_t)gf_context_stream(context())),
              "hipMemsetAsync");
  }
  HostVec<T> createHostVec() const {
    HostVec<T> h(N_);
    hip_check(hipMemcpyAsync(h.getPtr(), data_, sizeof(T) * N_,
                             hipMemcpyDeviceToHost,
                             (hipStream_t)gf_context_stream(context())),
              "D2H");
    synchronize();
    return h;
  }
  DeviceVec copyVec() const {
    DeviceVec v(N_, false);
    copyValuesTo(v);
    return v;
  }
  void copyValuesTo(DeviceVec dest) const {
    hip_check(hipMemcpyAsync(dest.getPtr(), data_, sizeof(T) * N_,
                             hipMemcpyDeviceToDevice,
                             (hipStream_t)gf_context_stream(context())),
              "D2D");
  }
  // perm is a device array (block_dim 6 only)
  void permuteData(int block_dim, const int* d_perm) {
    static_assert(std::is_same<T, double>::value, "FP64 vectors only");
    if (block_dim != 6) throw Error(GF_ERR_INVALID, "block_dim must be 6");
    DeviceVec tmp(N_, false);
    check(gf_permute_vec(context(), N_ / 6, d_perm, data_, tmp.getPtr()),
          "permute");
    tmp.copyValuesTo(*this);
  }
  DeviceVec createPermuteVec(int block_dim, const int* d_perm) const {
    DeviceVec v = copyVec();
    v.permuteData(block_dim, d_perm);
    return v;
  }
  void apply_bcs(const DeviceVec<int>& bcs) {
    check(gf_bc_apply_vec(context(), bcs.getSize(), bcs.getPtr(), nullptr,
                          (double*)data_),
          "apply_bcs(vec)");
  }
  DeviceVec createDeviceVec() const { return *this; }

 private:
  int N_ = 0;
  T* data_ = nullptr;
  std::shared_ptr<T> buf_;
};

template <typename T>
DeviceVec<T> HostVec<T>::createDeviceVec(bool memset) const {
  DeviceVec<T> d(N_, memset);
  hip_check(hipMemcpyAsync(d.getPtr(), data_, sizeof(T) * N_,
                           hipMemcpyHostToDevice,
                           (hipStream_t)gf_context_stream(context())),
            "H2D");
  synchronize();
  return d;
}

template <typename T>
using VecType = DeviceVec<T>;

// ----------------------------------------------------------- free functions
inline int TacsComputeRCMOrder(int nvars, const int* rowp, const int* cols,
                               int* rcm_vars, int root, int n_rcm_iters) {
  check(gf_rcm_order(nvars, rowp, cols, root, n_rcm_iters, rcm_vars), "RCM");
  return nvars;
}

inline int getBandWidth(const int& nnodes, const int& /*nnzb*/, int* rowPtr,
                        int* colPtr) {
  return gf_bandwidth(nnodes, rowPtr, colPtr);
}

namespace detail {
inline std::shared_ptr<int> own(int* p) {
  return std::shared_ptr<int>(p, [](int* q) { gf_free(q); });
}
inline std::shared_ptr<int> host_ints(const int* src, size_t n) {
  int* p = (int*)std::malloc(sizeof(int) * (n ? n : 1));
  if (n) std::memcpy(p, src, sizeof(int) * n);
  return own(p);
}
inline std::shared_ptr<int> device_ints(const int* src, size_t n) {
  int* p = nullptr;
  hip_check(hipMalloc(&p, sizeof(int) * (n ? n : 1)), "hipMalloc");
  if (n)
    hip_check(hipMemcpy(p, src, sizeof(int) * n, hipMemcpyHostToDevice), "H2D");
  return std::shared_ptr<int>(p, [](int* q) { hipFree(q); });
}
}  // namespace detail

// computeILUk (bsr_utils.h:678): replaces rowPtr/colPtr by the ILU(k) pattern;
// the returned arrays are owned by the caller (release with gf_free)
inline void computeILUk(int nnodes, int /*nnzb*/, int*& rowPtr, int*& colPtr,
                        int levFill, double /*fill*/, int** levs = nullptr) {
  int *rp = nullptr, *cp = nullptr, nnz = 0;
  check(gf_iluk_symbolic(nnodes, rowPtr, colPtr, levFill, &rp, &cp, &nnz),
        "computeILUk");
  rowPtr = rp;
  colPtr = cp;
  if (levs) *levs = nullptr;
}

// ------------------------------------------------------------------ BsrData
class BsrData {
 public:
  BsrData() = default;
  // node graph from connectivity (bsr_utils.h:46-62)
  BsrData(int nelems_, int nnodes_, const int& nodes_per_elem_,
          const int& block_dim_, const int32_t* conn_)
      : nelems(nelems_), nnodes(nnodes_), nodes_per_elem(nodes_per_elem_),
        block_dim(block_dim_), host(true) {
    conn_own_ = detail::host_ints(conn_, (size_t)nelems * nodes_per_elem);
    conn = conn_own_.get();
    int *rp = nullptr, *cp = nullptr;
    check(gf_node_graph(nelems, nnodes, nodes_per_elem, conn, &rp, &cp, &nnzb),
          "node graph");
    keep(rp, cp);
    make_nominal_ordering();
  }
  // pattern given (bsr_utils.h:64-82)
  BsrData(int nnodes_, int block_dim_, int nnzb_, int* rowPtr_, int* colPtr_,
          int* perm_ = nullptr, int* iperm_ = nullptr)
      : nnodes(nnodes_), block_dim(block_dim_), nnzb(nnzb_), host(true) {
    keep(detail::host_ints(rowPtr_, nnodes + 1),
         detail::host_ints(colPtr_, nnzb));
    if (perm_) {
      perm_own_ = detail::host_ints(perm_, nnodes);
      iperm_own_ = detail::host_ints(iperm_, nnodes);
      perm = perm_own_.get();
      iperm = iperm_own_.get();
    } else {
      make_nominal_ordering();
    }
  }

  // reordered pattern + perm[old] = new (bsr_utils.h:84-103)
  void post_reordering_update(int nnzb_, int* rowPtr_, int* colPtr_,
                              int* perm_, int* iperm_, bool /*print*/ = true) {
    nnzb = nnzb_;
    keep(detail::host_ints(rowPtr_, nnodes + 1), detail::host_ints(colPtr_, nnzb));
    perm_own_ = detail::host_ints(perm_, nnodes);
    iperm_own_ = detail::host_ints(iperm_, nnodes);
    perm = perm_own_.get();
    iperm = iperm_own_.get();
    update_maps();
  }

  // AMD reordering + complete symbolic fill (bsr_utils.h:105-145), the
  // pattern of the direct solve
  void symbolic_factorization(double /*fill_factor*/ = 10.0, bool /*print*/ = false) {
    std::vector<int> p(nnodes), ip(nnodes);
    check(gf_amd_order(nnodes, rowPtr, colPtr, p.data()), "AMD");
    for (int i = 0; i < nnodes; i++) ip[p[i]] = i;
    int *prp = nullptr, *pcp = nullptr, *frp = nullptr, *fcp = nullptr, nnz = 0;
    check(gf_pattern_permute(nnodes, rowPtr, colPtr, p.data(), &prp, &pcp),
          "permute");
    check(gf_iluk_symbolic(nnodes, prp, pcp, 1 << 29, &frp, &fcp, &nnz), "fill");
    gf_free(prp);
    gf_free(pcp);
    nnzb = nnz;
    keep(frp, fcp);
    perm_own_ = detail::host_ints(p.data(), nnodes);
    iperm_own_ = detail::host_ints(ip.data(), nnodes);
    perm = perm_own_.get();
    iperm = iperm_own_.get();
    update_maps();
  }

  // deep copy of every array onto the device (bsr_utils.h:213-280); the
  // device copy keeps the host arrays alive for host-side analyses
  BsrData createDeviceBsrData() const {
    BsrData d = *this;
    d.host = false;
    d.host_twin_ = std::make_shared<BsrData>(*this);
    d.rp_own_ = detail::device_ints(rowPtr, nnodes + 1);
    d.cp_own_ = detail::device_ints(colPtr, nnzb);
    d.rowPtr = d.rp_own_.get();
    d.colPtr = d.cp_own_.get();
    d.perm_own_ = detail::device_ints(perm, nnodes);
    d.iperm_own_ = detail::device_ints(iperm, nnodes);
    d.perm = d.perm_own_.get();
    d.iperm = d.iperm_own_.get();
    if (elemIndMap) {
      d.emap_own_ = detail::device_ints(elemIndMap, (size_t)nelems * nodes_per_elem *
                                                        nodes_per_elem);
      d.elemIndMap = d.emap_own_.get();
    }
    if (transpose_rowPtr) {
      d.tr_own_ = detail::device_ints(transpose_rowPtr, nnodes + 1);
      d.tc_own_ = detail::device_ints(transpose_colPtr, nnzb);
      d.tm_own_ = detail::device_ints(transpose_block_map, nnzb);
      d.transpose_rowPtr = d.tr_own_.get();
      d.transpose_colPtr = d.tc_own_.get();
      d.transpose_block_map = d.tm_own_.get();
    }
    d.conn = nullptr;
    return d;
  }
  const BsrData& hostData() const { return host ? *this : *host_twin_; }
  int getNumValues() const { return block_dim * block_dim * nnzb; }

  int nelems = 0, nnodes = 0, nodes_per_elem = 0, block_dim = 0, nnzb = 0;
  const int32_t* conn = nullptr;
  int *rowPtr = nullptr, *colPtr = nullptr, *elemIndMap = nullptr;
  int *perm = nullptr, *iperm = nullptr;
  int *transpose_rowPtr = nullptr, *transpose_colPtr = nullptr,
      *transpose_block_map = nullptr;
  bool host = true;

 private:
  void keep(int* rp, int* cp) { keep(detail::own(rp), detail::own(cp)); }
  void keep(std::shared_ptr<int> rp, std::shared_ptr<int> cp) {
    rp_own_ = rp;
    cp_own_ = cp;
    rowPtr = rp.get();
    colPtr = cp.get();
  }
  void make_nominal_ordering() {
    std::vector<int> p(nnodes);
    for (int i = 0; i < nnodes; i++) p[i] = i;
    perm_own_ = detail::host_ints(p.data(), nnodes);
    iperm_own_ = detail::host_ints(p.data(), nnodes);
    perm = perm_own_.get();
    iperm = iperm_own_.get();
  }
  void update_maps() {
    if (conn) {
      std::vector<int> m((size_t)nelems * nodes_per_elem * nodes_per_elem);
      check(gf_elem_ind_map(nelems, nodes_per_elem, conn, rowPtr, colPtr, perm,
                            m.data()),
            "elem_ind_map");
      emap_own_ = detail::host_ints(m.data(), m.size());
      elemIndMap = emap_own_.get();
    }
    int *tr = nullptr, *tc = nullptr, *tm = nullptr;
    check(gf_transpose_map(nnodes, rowPtr, colPtr, &tr, &tc, &tm), "transpose");
    tr_own_ = detail::own(tr);
    tc_own_ = detail::own(tc);
    tm_own_ = detail::own(tm);
    transpose_rowPtr = tr;
    transpose_colPtr = tc;
    transpose_block_map = tm;
  }
  std::shared_ptr<int> conn_own_, rp_own_, cp_own_, perm_own_, iperm_own_,
      emap_own_, tr_own_, tc_own_, tm_own_;
  std::shared_ptr<BsrData> host_twin_;
};

// ------------------------------------------------------------------- BsrMat
template <class Vec_>
class BsrMat {
 public:
  using T = typename Vec_::type;
  using Vec = Vec_;
  BsrMat(const BsrData& bsr_data, Vec& values)
      : bsr_data_(bsr_data), values_(values) {}
  explicit BsrMat(const BsrData& bsr_data)
      : bsr_data_(bsr_data), values_(bsr_data.getNumValues()) {}
  void zeroValues() { values_.zeroValues(); }
  BsrData getBsrData() const { return bsr_data_; }
  int get_nnz() const { return bsr_data_.getNumValues(); }
  Vec getVec() { return values_; }
  HostVec<T> createHostVec() { return values_.createHostVec(); }
  T* getPtr() { return values_.getPtr(); }
  int* getPerm() { return bsr_data_.perm; }
  int* getIPerm() { return bsr_data_.iperm; }
  int getBlockDim() const { return bsr_data_.block_dim; }
  int* getRowPtr() { return bsr_data_.rowPtr; }
  int* getColPtr() { return bsr_data_.colPtr; }
  // rows and columns of constrained DOFs zeroed, unit diagonal
  void apply_bcs(const DeviceVec<int>& bcs) {
    check(gf_bc_apply_mat(context(), bcs.getSize(), bcs.getPtr(),
                          bsr_data_.rowPtr, bsr_data_.colPtr,
                          bsr_data_.transpose_rowPtr,
                          bsr_data_.transpose_colPtr,
                          bsr_data_.transpose_block_map, bsr_data_.perm,
                          values_.getPtr()),
          "apply_bcs(mat)");
  }
  // copy into a matrix whose pattern is a superset (preconditioner)
  void copyValuesTo(BsrMat mat) {
    check(gf_bsr_copy_to_pattern(context(), bsr_data_.nnodes, bsr_data_.rowPtr,
                                 bsr_data_.colPtr, values_.getPtr(),
                                 mat.getRowPtr(), mat.getColPtr(), mat.getPtr()),
          "copyValuesTo");
  }

 private:
  BsrData bsr_data_;
  Vec values_;
};

template <class Assembler, class Vec>
BsrMat<Vec> createBsrMat(Assembler& assembler) {
  return BsrMat<Vec>(assembler.getBsrData());
}

// ------------------------------------------------------ ILU preconditioner
template <class Vec_>
class ILUk_Preconditioner {
 public:
  using Vec = Vec_;
  using T = typename Vec::type;
  ILUk_Preconditioner(BsrMat<Vec>& precond_mat, BsrMat<Vec>& orig_mat,
                      bool /*print*/ = false)
      : mat_(orig_mat), precond_mat_(precond_mat) {
    const BsrData& h = precond_mat_.getBsrData().hostData();
    gf_ilu p = nullptr;
    check(gf_ilu_create(context(), h.nnodes, h.rowPtr, h.colPtr, &p),
          "ilu analysis");
    ilu_ = std::shared_ptr<gf_ilu_s>(p, gf_ilu_destroy);
  }
  void copyValues() { mat_.copyValuesTo(precond_mat_); }
  // numeric ILU in FP64 on the preconditioner pattern
  void factor() {
    copyValues();
    int zp = -1;
    check(gf_bsr_ilu_factor_f64(ilu_.get(), precond_mat_.getPtr(), &zp),
          "ILU factor");
  }
  // x = U^-1 L^-1 b
  void solve(const Vec& b, Vec& x) {
    Vec tmp(b.getSize(), false);
    check(gf_bsr_ilu_apply_f64(ilu_.get(), precond_mat_.getPtr(), b.getPtr(),
                               tmp.getPtr(), x.getPtr()),
          "ILU apply");
  }
  gf_ilu handle() { return ilu_.get(); }
  BsrMat<Vec>& matrix() { return mat_; }
  BsrMat<Vec>& precond_matrix() { return precond_mat_; }

 private:
  BsrMat<Vec>& mat_;
  BsrMat<Vec>& precond_mat_;
  std::shared_ptr<gf_ilu_s> ilu_;
};

// ------------------------------------------------------- solve utilities
template <class Vec>
Vec bsr_pre_solve(BsrMat<Vec> mat, Vec rhs, Vec /*soln*/) {
  return rhs.createPermuteVec(mat.getBlockDim(), mat.getPerm());
}
template <class Vec>
void bsr_post_solve(BsrMat<Vec> mat, Vec /*rhs*/, Vec soln) {
  soln.permuteData(mat.getBlockDim(), mat.getIPerm());
}

namespace solvers {
struct GmresStats {
  int iterations = 0;
  double rel_resid = 0.0;
};
// Right-preconditioned restarted GMRES(m) in FP64 on the permuted system
// (rhs and soln in the matrix numbering, e.g. from bsr_pre_solve).
template <class Vec>
GmresStats gmres(BsrMat<Vec>& kmat, ILUk_Preconditioner<Vec>& precond,
                 Vec& rhs, Vec& soln, int restart = 50, int max_iters = 5000,
                 double rtol = 1e-8) {
  GmresStats s;
  const BsrData d = kmat.getBsrData();
  check(gf_gmres_solve(context(), d.nnodes, d.rowPtr, d.colPtr, kmat.getPtr(),
                       nullptr, precond.handle(),
                       precond.precond_matrix().getPtr(), nullptr,
                       rhs.getPtr(), soln.getPtr(), restart, max_iters, rtol, 0,
                       &s.iterations, &s.rel_resid),
        "gmres");
  return s;
}
// Direct solve: the matrix must be on its complete fill pattern
// (BsrData::symbolic_factorization); ILU there is the exact block LU.
template <class Vec>
void direct_LU_solve(BsrMat<Vec>& mat, Vec& rhs, Vec& soln) {
  const BsrData& h = mat.getBsrData().hostData();
  gf_ilu p = nullptr;
  check(gf_ilu_create(context(), h.nnodes, h.rowPtr, h.colPtr, &p), "LU analysis");
  std::shared_ptr<gf_ilu_s> ilu(p, gf_ilu_destroy);
  Vec lu = mat.getVec().copyVec();
  int zp = -1;
  check(gf_bsr_ilu_factor_f64(ilu.get(), lu.getPtr(), &zp), "LU factor");
  Vec b = bsr_pre_solve(mat, rhs, soln);
  Vec tmp(b.getSize(), false);
  check(gf_bsr_ilu_apply_f64(ilu.get(), lu.getPtr(), b.getPtr(), tmp.getPtr(),
                             soln.getPtr()),
        "LU solve");
  check(gf_ilu_check(ilu.get()), "LU solve");
  bsr_post_solve(mat, rhs, soln);
}
// max-magnitude entry (signed) of K u - f, as CUSPARSE::get_resid
// (include/linalg/solvers/cusparse.h:853-918); rhs and soln in the matrix
// numbering
template <class Vec>
typename Vec::type get_resid(BsrMat<Vec> mat, Vec rhs, Vec soln) {
  const BsrData d = mat.getBsrData();
  double mx = 0.0, nrm = 0.0;
  check(gf_bsr_resid_f64(context(), d.nnodes, d.rowPtr, d.colPtr, mat.getPtr(),
                         soln.getPtr(), rhs.getPtr(), &mx, &nrm),
        "get_resid");
  return mx;
}
}  // namespace solvers

// ------------------------------------------------------- shell type tags
template <typename T>
struct QuadLinearQuadrature {
  static constexpr int num_quad_pts = 4;
};
template <typename T, int offset = 3>
struct LinearizedRotation {};
template <typename T, class Quadrature_, int order = 2>
struct ShellQuadBasis {
  static_assert(order == 2, "MITC4 (order 2) only");
  using Quadrature = Quadrature_;
  static constexpr int num_nodes = 4;
  struct LinearQuadGeo {
    static constexpr int spatial_dim = 3;
    static constexpr int num_nodes = 4;
  };
  using Geo = LinearQuadGeo;
};
template <typename T, bool has_ref_axis_ = false>
struct ShellIsotropicData {
  static_assert(!has_ref_axis_, "reference-axis transform is not supported");
  static constexpr bool has_ref_axis = has_ref_axis_;
  ShellIsotropicData() = default;
  ShellIsotropicData(T E_, T nu_, T thick_, T tOffset_ = 0.0)
      : E(E_), nu(nu_), thick(thick_), tOffset(tOffset_) {}
  T E = 0, nu = 0, thick = 0, tOffset = 0;
};
// include/shell/physics.h:6-17: isNonlinear selects the strain of
// NonlinearShellStrainCore (a2d/a2dshellstrain.h:72-93); the assembler then
// forms the exact tangent and residual at the current variables
template <typename T, class Data_, bool isNonlinear = false>
struct IsotropicShell {
  using Data = Data_;
  static constexpr int vars_per_node = 6;
  static constexpr bool is_nonlinear = isNonlinear;
};
template <typename T, class Director_, class Basis_, class Phys_>
struct ShellElementGroup {
  using Director = Director_;
  using Basis = Basis_;
  using Geo = typename Basis::Geo;
  using Phys = Phys_;
};

// ------------------------------------------------------- TACSMeshLoader
// NASTRAN bulk-data reader (include/mesh/TACSMeshLoader.h:145-227) on
// gf_bdf_read.  scanBDFFile returns 0 or 1 as the reference does (the reason
// is in error()); getAssemblerCreatorData hands out new[] arrays the caller
// deletes, as the reference does.
template <typename T>
class TACSMeshLoader {
 public:
  TACSMeshLoader() = default;
  TACSMeshLoader(const TACSMeshLoader&) = delete;
  TACSMeshLoader& operator=(const TACSMeshLoader&) = delete;
  ~TACSMeshLoader() { gf_bdf_destroy(h_); }

  int scanBDFFile(const char* file_name) {
    gf_bdf_destroy(h_);
    h_ = nullptr;
    if (gf_bdf_read(file_name, &h_) != GF_OK) {
      error_ = gf_bdf_error();
      return 1;
    }
    gf_bdf_sizes(h_, &num_nodes_, &num_elements_, &conn_size_, nullptr,
                 &num_bc_vars_, &num_components_);
    return 0;
  }
  const std::string& error() const { return error_; }
  // extensions: total element nodes, and whether every SPC component digit
  // names one of nv DOFs (digits 7-8 would spill into the next node)
  int getElemConnSize() { return conn_size_; }
  bool bcVarsWithin(int nv) {
    std::vector<int> v(num_bc_vars_ > 0 ? num_bc_vars_ : 1);
    gf_bdf_bcs(h_, nullptr, nullptr, v.data(), nullptr);
    for (int j = 0; j < num_bc_vars_; j++)
      if (v[j] >= nv) return false;
    return true;
  }
  int getNumNodes() { return num_nodes_; }
  int getNumElements() { return num_elements_; }
  int getNumComponents() { return num_components_; }
  const char* getComponentDescript(int comp) {
    return gf_bdf_component_descript(h_, comp);
  }
  const char* getElementDescript(int comp) {
    return gf_bdf_element_descript(h_, comp);
  }
  void getAssemblerCreatorData(const int vars_per_node, int& my_num_nodes,
                               int& my_num_elements, int& my_num_bcs,
                               int*& my_elem_conn, int*& my_bcs,
                               T*& my_xpts) {
    if (!h_) throw Error(GF_ERR_INVALID, "getAssemblerCreatorData: no mesh");
    my_num_nodes = num_nodes_;
    my_num_elements = num_elements_;
    my_num_bcs = num_bc_vars_;
    my_elem_conn = new int[conn_size_ > 0 ? conn_size_ : 1];
    my_bcs = new int[num_bc_vars_ > 0 ? num_bc_vars_ : 1];
    std::vector<double> x(3 * (size_t)num_nodes_);
    gf_bdf_assembler_data(h_, vars_per_node, my_elem_conn, my_bcs, x.data());
    my_xpts = new T[3 * (size_t)num_nodes_ + 1];
    for (size_t i = 0; i < x.size(); i++) my_xpts[i] = T(x[i]);
  }

 private:
  gf_bdf h_ = nullptr;
  std::string error_;
  int num_nodes_ = 0, num_elements_ = 0, conn_size_ = 0, num_bc_vars_ = 0,
      num_components_ = 0;
};

// --------------------------------------------------------- ElementAssembler
template <typename T_, typename ElemGroup, template <typename> class Vec,
          template <typename> class Mat_>
class ElementAssembler {
 public:
  using T = T_;
  using Geo = typename ElemGroup::Geo;
  using Basis = typename ElemGroup::Basis;
  using Phys = typename ElemGroup::Phys;
  using Data = typename Phys::Data;
  using Mat = Mat_<Vec<T>>;
  static_assert(std::is_same<T, double>::value, "FP64 assembly only");
  static constexpr int32_t vars_per_node = Phys::vars_per_node;

  ElementAssembler(int32_t num_geo_nodes_, int32_t num_vars_nodes_,
                   int32_t num_elements_, HostVec<int32_t>& geo_conn,
                   HostVec<int32_t>& vars_conn, HostVec<T>& xpts,
                   HostVec<int>& bcs, HostVec<Data>& physData)
      : num_geo_nodes(num_geo_nodes_), num_vars_nodes(num_vars_nodes_),
        num_elements(num_elements_) {
    if (num_geo_nodes != num_vars_nodes ||
        std::memcmp(geo_conn.getPtr(), vars_conn.getPtr(),
                    sizeof(int) * 4 * num_elements) != 0)
      throw Error(GF_ERR_INVALID, "geometry and variable meshes must coincide");
    vars_ = Vec<T>(get_num_vars());
    bsr_data = BsrData(num_elements, num_vars_nodes, 4, 6, vars_conn.getPtr());
    conn_ = vars_conn.createDeviceVec();
    xpts_ = xpts.createDeviceVec();
    bcs_ = bcs.createDeviceVec();
    HostVec<T> d(4 * num_elements);
    for (int e = 0; e < num_elements; e++) {
      d[4 * e] = physData[e].E;
      d[4 * e + 1] = physData[e].nu;
      d[4 * e + 2] = physData[e].thick;
      d[4 * e + 3] = physData[e].tOffset;
    }
    data_ = d.createDeviceVec();
  }

  // include/assembler.h:92-115: one Data for every element; every element
  // must be a 4-node quad (CQUAD4 / CQUADR)
  static ElementAssembler createFromBDF(TACSMeshLoader<T>& mesh_loader,
                                        Data single_data) {
    int num_nodes = 0, num_elements = 0, num_bcs = 0;
    int *elem_conn = nullptr, *bcs = nullptr;
    T* xpts = nullptr;
    mesh_loader.getAssemblerCreatorData(vars_per_node, num_nodes, num_elements,
                                        num_bcs, elem_conn, bcs, xpts);
    std::unique_ptr<int[]> c(elem_conn), b(bcs);
    std::unique_ptr<T[]> x(xpts);
    if (mesh_loader.getElemConnSize() != 4 * num_elements)
      throw Error(GF_ERR_INVALID, "createFromBDF: every element must be a "
                                  "4-node quad (CQUAD4 / CQUADR)");
    if (!mesh_loader.bcVarsWithin(vars_per_node))
      throw Error(GF_ERR_INVALID, "createFromBDF: SPC component digit beyond "
                                  "the node's DOFs");
    HostVec<int32_t> conn_vec(4 * num_elements, elem_conn);
    HostVec<int> bcs_vec(num_bcs, bcs);
    HostVec<T> xpts_vec(3 * num_nodes, xpts);
    HostVec<Data> phys(num_elements, single_data);
    return ElementAssembler(num_nodes, num_nodes, num_elements, conn_vec,
                            conn_vec, xpts_vec, bcs_vec, phys);
  }

  BsrData getBsrData() { return bsr_data; }
  void symbolic_factorization(double fillin = 100.0, bool print = false) {
    bsr_data.symbolic_factorization(fillin, print);
  }
  int get_num_xpts() { return num_geo_nodes * 3; }
  int get_num_vars() { return num_vars_nodes * vars_per_node; }
  int get_num_nodes() { return num_vars_nodes; }
  int get_num_elements() { return num_elements; }
  Vec<T> getXpts() { return xpts_; }
  Vec<int> getConn() { return conn_; }

  void apply_bcs(Vec<T>& vec, bool = false) { vec.apply_bcs(bcs_); }
  void apply_bcs(Mat& mat, bool = false) { mat.apply_bcs(bcs_); }

  Vec<T> createVarsVec(T* data = nullptr, bool randomize = false) {
    HostVec<T> h(get_num_vars());
    if (data) std::memcpy(h.getPtr(), data, sizeof(T) * get_num_vars());
    if (randomize) h.randomize();
    return h.createDeviceVec();
  }
  // variables in the unpermuted node numbering (the device gather uses the
  // mesh connectivity directly)
  void set_variables(Vec<T>& newVars) { newVars.copyValuesTo(vars_); }

  void add_energy(T& Uenergy, bool = false) {
    DeviceVec<T> e(1);
    if constexpr (Phys::is_nonlinear)
      check(gf_assemble_tangent_mitc4_f64(
                context(), num_elements, num_vars_nodes, conn_.getPtr(),
                xpts_.getPtr(), vars_.getPtr(), data_.getPtr(), nullptr,
                nullptr, nullptr, nullptr, e.getPtr()),
            "add_energy");
    else
      check(gf_energy_mitc4_f64(context(), num_elements, conn_.getPtr(),
                                xpts_.getPtr(), vars_.getPtr(), data_.getPtr(),
                                e.getPtr()),
            "add_energy");
    Uenergy = e.createHostVec()[0];
  }
  // residual accumulated at 6*perm[node]+k of the matrix numbering
  void add_residual(Vec<T>& res, bool = false) {
    const int* perm = device_perm();
    if constexpr (Phys::is_nonlinear)
      check(gf_assemble_tangent_mitc4_f64(
                context(), num_elements, num_vars_nodes, conn_.getPtr(),
                xpts_.getPtr(), vars_.getPtr(), data_.getPtr(), nullptr, perm,
                nullptr, res.getPtr(), nullptr),
            "add_residual");
    else
      check(gf_assemble_residual_mitc4_f64(context(), num_elements,
                                           num_vars_nodes, conn_.getPtr(),
                                           xpts_.getPtr(), vars_.getPtr(),
                                           data_.getPtr(), perm, res.getPtr()),
            "add_residual");
  }
  // K accumulated into mat (zeroed here), residual into res; the tangent
  // K(u) and R(u) at the current variables for the nonlinear element
  void add_jacobian(Vec<T>& res, Mat& mat, bool = false) {
    BsrData d = mat.getBsrData();
    if (d.host || !d.elemIndMap)
      throw Error(GF_ERR_INVALID,
                  "add_jacobian: matrix needs device BsrData with elemIndMap");
    mat.zeroValues();
    if constexpr (Phys::is_nonlinear)
      check(gf_assemble_tangent_mitc4_f64(
                context(), num_elements, num_vars_nodes, conn_.getPtr(),
                xpts_.getPtr(), vars_.getPtr(), data_.getPtr(), d.elemIndMap,
                d.perm, mat.getPtr(), res.getPtr(), nullptr),
            "add_jacobian");
    else
      check(gf_assemble_jacobian_mitc4_f64(
                context(), num_elements, num_vars_nodes, conn_.getPtr(),
                xpts_.getPtr(), vars_.getPtr(), data_.getPtr(), d.elemIndMap,
                d.perm, mat.getPtr(), res.getPtr()),
            "add_jacobian");
    perm_cache_ = d;
  }

  BsrData bsr_data;

 private:
  const int* device_perm() {
    if (!perm_cache_.host && perm_cache_.perm) return perm_cache_.perm;
    perm_cache_ = bsr_data.createDeviceBsrData();
    return perm_cache_.perm;
  }
  int32_t num_geo_nodes, num_vars_nodes, num_elements;
  Vec<int> conn_, bcs_;
  Vec<T> xpts_, vars_, data_;
  BsrData perm_cache_;
};

// ------------------------------------------------------ plate generator
// examples/plate_demo/_plate_utils.h:9-144 with the connectivity order as an
// option: tensor order by default (the basis' node order), counter-clockwise
// (the reference generator's order) when ccw = true.
template <class Assembler>
Assembler createPlateAssembler(int nxe, int nye, double Lx, double Ly, double E,
                               double nu, double thick, bool ccw = false) {
  using T = typename Assembler::T;
  using Data = typename Assembler::Data;
  const int nnx = nxe + 1, nny = nye + 1, num_nodes = nnx * nny;
  const int num_elements = nxe * nye;
  std::vector<int> b;
  for (int k = 0; k < 6; k++) b.push_back(k);
  for (int ix = 1; ix < nnx; ix++) {
    b.push_back(6 * ix + 1);
    b.push_back(6 * ix + 2);
  }
  for (int iy = 1; iy < nny; iy++) {
    int n = nnx * iy;
    b.push_back(6 * n);
    b.push_back(6 * n + 2);
    n = nnx * iy + nnx - 1;
    b.push_back(6 * n + 2);
  }
  for (int ix = 1; ix < nnx - 1; ix++) b.push_back(6 * (nnx * (nny - 1) + ix) + 2);
  HostVec<int> bcs((int)b.size());
  for (size_t i = 0; i < b.size(); i++) bcs[i] = b[i];
  HostVec<int32_t> conn(4 * num_elements);
  for (int iy = 0; iy < nye; iy++)
    for (int ix = 0; ix < nxe; ix++) {
      const int e = nxe * iy + ix, n0 = nnx * iy + ix;
      int* c = &conn[4 * e];
      c[0] = n0;
      c[1] = n0 + 1;
      c[2] = ccw ? n0 + nnx + 1 : n0 + nnx;
      c[3] = ccw ? n0 + nnx : n0 + nnx + 1;
    }
  HostVec<T> xpts(3 * num_nodes);
  const T dx = Lx / nxe, dy = Ly / nye;
  for (int iy = 0; iy < nny; iy++)
    for (int ix = 0; ix < nnx; ix++) {
      T* x = &xpts[3 * (nnx * iy + ix)];
      x[0] = dx * ix;
      x[1] = dy * iy;
      x[2] = 0.0;
    }
  HostVec<Data> physData(num_elements, Data(E, nu, thick));
  return Assembler(num_nodes, num_nodes, num_elements, conn, conn, xpts, bcs,
                   physData);
}

// ------------------------------------------------------------ printToVTK
// include/mesh/vtk_writer.h:8-73: mesh and displacement field (soln in the
// unpermuted node order, host or device vector) as legacy VTK
namespace detail {
template <typename T>
HostVec<T> to_host(const HostVec<T>& v) { return v; }
template <typename T>
HostVec<T> to_host(const DeviceVec<T>& v) { return v.createHostVec(); }
}  // namespace detail

template <class Assembler, class Vec>
void printToVTK(Assembler assembler, Vec soln, std::string filename) {
  auto x = assembler.getXpts().createHostVec();
  auto c = assembler.getConn().createHostVec();
  auto u = detail::to_host(soln);
  if (u.getSize() != assembler.get_num_vars())
    throw Error(GF_ERR_INVALID, "printToVTK: solution size");
  check(gf_write_vtk(filename.c_str(), assembler.get_num_nodes(), x.getPtr(),
                     assembler.get_num_elements(), c.getPtr(), u.getPtr()),
        "printToVTK");
}

// point-sampled sine load on w (examples/plate_demo/_plate_utils.h:146-184)
template <typename T, class Phys>
std::vector<T> getPlateLoads(int nxe, int nye, double Lx, double Ly,
                             double load_mag) {
  const int nnx = nxe + 1, nny = nye + 1;
  const double PI = 3.1415926535897;  // the reference's literal
  std::vector<T> f(6 * nnx * nny, T(0));
  for (int iy = 0; iy < nny; iy++)
    for (int ix = 0; ix < nnx; ix++) {
      const T x = ix * Lx / nxe, y = iy * Ly / nye;
      f[6 * (nnx * iy + ix) + 2] =
          load_mag * std::sin(PI * x / Lx) * std::sin(PI * y / Ly);
    }
  return f;
}

}  // namespace gpufem
