// GMRES driver interfaces (see krylov.hip).
#pragma once
#include <cstddef>
#include <vector>

#include "common.h"

namespace gf {

// Device workspace for GMRES(m) on N = 6 * nnodes unknowns.
struct KrylovWork {
  gf_context ctx = nullptr;
  size_t N = 0;
  int m = 0;
  double *V = nullptr, *w = nullptr, *z = nullptr, *r = nullptr,
         *tmp = nullptr;
  double *h = nullptr, *h2 = nullptr, *nrm = nullptr, *y = nullptr;
  double *state = nullptr, *partial = nullptr, *scal = nullptr;
  int* gate = nullptr;  // second Gram-Schmidt pass needed (device flag)
  // DCGS2: Hessenberg before rotations (column-major, (m+1) x (m+1)), the
  // step's dot products, update coefficients, provisional subdiagonal
  double *hc = nullptr, *dd = nullptr, *co = nullptr, *rho = nullptr;
  float *a32 = nullptr, *b32 = nullptr, *c32 = nullptr;
  double* host_scal = nullptr;  // pinned: two (estimate, h_{j+1,j}) slots
  // Z_i = M^-1 V_i kept beside the basis (m vectors; null when not kept):
  // the cycle's update is then x += Z y, without a preconditioner
  // application
  double* Z = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int init(gf_context c, size_t N, int m, bool mixed, bool keep_z = false);
  void release();
  // out[0..nvec) = V_i . x (mode 0 store, 1 add, 2 sqrt); skipped when
  // *gate == 0 (a stored result then reads zero)
  int dots(const double* V, size_t ldv, int nvec, const double* x, double* out,
           int mode, const int* gate = nullptr);
};

// The linear operator A (BSR, FP64 and optional FP32 copy) and the ILU
// preconditioner M (factored values in FP64 and/or FP32).
struct Operator {
  gf_context ctx = nullptr;
  int n = 0;
  const int *krowp = nullptr, *kcols = nullptr;
  const double* kvals = nullptr;
  const float* kvals32 = nullptr;
  gf_ilu ilu = nullptr;
  const double* mlu = nullptr;
  const float* mlu32 = nullptr;
  double *work64 = nullptr, *work64b = nullptr;
  float *t32a = nullptr, *t32b = nullptr, *t32c = nullptr;
  int apply_a64(const double* in, double* out) const;
  // prec: 0 FP64 ILU; 1 FP32 ILU and FP32 SpMV (inner loop of GMRES-IR);
  // 2 FP32 ILU, FP64 SpMV
  int apply_prec(const double* in, double* out, int prec) const;
  // out = A M^-1 in; z_out (may be null, precisions 0 and 2) receives
  // M^-1 in in FP64
  int apply_prec_a(const double* in, double* out, int prec,
                   double* z_out = nullptr) const;
};

struct GmresResult {
  int iterations = 0;
  int outer = 0;
  double rel_resid = 0.0;
};

// prec as in Operator: 1 runs GMRES-IR (inner cycles stop at a 1e-5
// reduction of the FP64 residual), 0 and 2 run plain restarted GMRES.
// stall_cycles > 0: stop (GF_ERR_NO_CONVERGENCE) once the true residual has
// fallen by less than 10% over that many restart cycles
int gmres_solve(KrylovWork& W, const Operator& op, const double* b, double* x,
                int max_iters, int stall_cycles, double rtol, int prec,
                GmresResult* res);

// The sweeps' wait-expired flag (device int, set by a triangular sweep whose
// bounded dependency wait ran out) and the context an ILU handle launches on.
int* ilu_sweep_err(gf_ilu h);
gf_context ilu_context(gf_ilu h);

// ---- batched sweeps (several systems in one launch per sweep)
// The arrays of one system in a batched sweep launch: L sweep b -> tmp
// (setting x to the sentinel row by row), U sweep tmp -> x.  tmp must hold
// the sentinel before the L sweep (see fill in the batched GMRES step).
template <typename T>
struct SweepSlot {
  const int4* lmeta;
  const int4* umeta;
  const int* lchunk;
  const int* uchunk;
  const int* cols;
  const T* lu;
  const T* b;
  T* tmp;
  T* x;
  const int* go;  // batched launches sweep the system only while *go == 1
};
// The device schedule of an ILU analysis (chunk mode only) and the
// dependency level of each chunk.
struct IluView {
  int n = 0;
  const int4 *lmeta = nullptr, *umeta = nullptr;
  const int *lchunk = nullptr, *uchunk = nullptr, *cols = nullptr;
  const std::vector<int>* lchunk_level = nullptr;
  const std::vector<int>* uchunk_level = nullptr;
  int grid = 0;
};
int ilu_view(gf_ilu h, IluView* v);
template <typename T>
int launch_batch_sweeps(hipStream_t s, const SweepSlot<T>* slots,
                        const int2* lmerged, int nl, const int2* umerged,
                        int nu, int grid, unsigned long long* counters,
                        int* err);
// bit pattern the sweeps take for "not yet written"
unsigned long long sweep_sentinel_bits(bool fp64);

// ---- batched GMRES of the serving loop
// Solves of several assembled and factored problems advance together: one
// Arnoldi step of every solve is one fused launch sequence (gather and
// precondition rhs, batched L and U sweeps, batched SpMV, batched DCGS2 dot
// products, coefficients and basis update), and a fixed number of steps is
// captured once as a HIP graph and replayed until the set of solves
// changes.  Right-preconditioned restarted GMRES as gmres_solve (DCGS2,
// cycle target 0.3 rtol, compensated FP64 residual at every cycle start),
// with the preconditioner in FP64 (precision 0) or FP32 (precision 2).
struct BatchJob {
  gf_problem p = nullptr;      // assembled and factored
  const double* loads = nullptr;  // host, 6 n, unpermuted
  double* x = nullptr;            // host, 6 n, unpermuted
  gf_solve_info info{};
  int status = GF_OK;
  bool done = false;
};
class BatchEngine;
// device: the device of the problems; precision 0 or 2
int batch_engine_create(int device, int precision, int max_slots,
                        BatchEngine** out);
void batch_engine_destroy(BatchEngine* e);
// queues a job and waits until its solve has finished (thread-safe)
int batch_engine_solve(BatchEngine* e, BatchJob* job);

}  // namespace gf
