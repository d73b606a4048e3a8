// C ABI of the engine (declarations and reference citations in
// include/pysa_amd.h).  Validation of user arguments and the reference's
// error messages live in the Python host layer; this layer checks what the
// kernels rely on (sizes, pointers) and maps HIP failures to codes.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pysa_amd.h"
#include "blocked.hpp"
#include "engine.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail(PYSA_AMD_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b ? b : 1);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct Problem {
  int n = 0, ptype = 0, dtype = 0, device = 0;
  bool symmetric = true;
  void* dJ = nullptr;   // row-major J
  void* dJt = nullptr;  // J transposed (== dJ when J is exactly symmetric)
  void* dJp = nullptr;  // rows of dJt padded with zeros to ldp (== dJt when ldp == n)
  int ldp = 0;  // a multiple of 256
  void* dh = nullptr;
  // sparse copy of a symmetric J with at most n/4 nonzeros per row: row i's
  // nonzeros in ascending column order, padded to ell_d with column -1
  int* d_ell_cols = nullptr;
  void* d_ell_vals = nullptr;
  int ell_d = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // calls on one problem are serialised
  // scratch reused across calls
  DevBuf fields, best_rep, best_rep_e, best_rep_sw, work;
  pysa::BlockedProblem blocked;
  ~Problem() {
    (void)hipSetDevice(device);
    pysa::blocked_release(blocked);
    if (dJp && dJp != dJt) (void)hipFree(dJp);
    if (dJt && dJt != dJ) (void)hipFree(dJt);
    if (dJ) (void)hipFree(dJ);
    if (dh) (void)hipFree(dh);
    if (d_ell_cols) (void)hipFree(d_ell_cols);
    if (d_ell_vals) (void)hipFree(d_ell_vals);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

Problem* get(int64_t h) { return reinterpret_cast<Problem*>(h); }

size_t esize(int dtype) { return dtype == 0 ? 4 : 8; }

// Host-side transpose (J may be only approximately symmetric; the reference
// always reads row `pos`, so the field update of column p needs J[:, p]).
template <typename T>
bool transpose_if_needed(const T* J, int n, std::vector<T>& out) {
  bool sym = true;
  for (int i = 0; i < n && sym; ++i)
    for (int j = i + 1; j < n; ++j)
      if (J[(size_t)i * n + j] != J[(size_t)j * n + i]) {
        sym = false;
        break;
      }
  if (sym) return true;
  out.resize((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[(size_t)j * n + i] = J[(size_t)i * n + j];
  return false;
}

struct RunArgs {
  int reads, R, S, order, use_pt, logw, best_sem, form, block_b;
  uint64_t seed, sweep_offset;
  const void* betas;
  const int32_t* bsort;
  const double* rep_lr;
  const int32_t* rep_perm;
  const double* rep_pt;
  int8_t* states;
  void* energies;
  int64_t* beta_idx;
  int8_t* best_state;
  void* best_energy;
  int32_t* best_sweeps;
  double* log_omegas;
  uint64_t* accepted;
};

__global__ void arange_i64(int64_t* bi, int reads, int R) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (long long)reads * R) bi[i] = i % R;
}
template <typename T>
__global__ void init_best(const T* energies, T* best_rep_e, int* best_rep_sw, long long count) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    best_rep_e[i] = energies[i];
    best_rep_sw[i] = 0;
  }
}

int check_run(const Problem* P, const RunArgs& a) {
  if (!P) return fail(PYSA_AMD_EINVAL, "invalid problem handle");
  if (a.reads < 1 || a.R < 1 || a.S < 1) return fail(PYSA_AMD_EINVAL, "reads, R and S must be >= 1");
  if (a.R > 1024) return fail(PYSA_AMD_ENOTSUP, "at most 1024 replicas per read");
  if (!a.betas || !a.states || !a.energies || !a.beta_idx || !a.best_state || !a.best_energy ||
      !a.best_sweeps)
    return fail(PYSA_AMD_EINVAL, "null array argument");
  if (a.logw && (!a.bsort || !a.log_omegas))
    return fail(PYSA_AMD_EINVAL, "log_omega needs beta_sort and log_omegas");
  if (a.order < 0 || a.order > 3) return fail(PYSA_AMD_EINVAL, "order must be 0, 1, 2 or 3");
  if (a.best_sem < 0 || a.best_sem > 1) return fail(PYSA_AMD_EINVAL, "best_semantics must be 0 or 1");
  if (a.block_b != 0 && a.block_b != 32)
    return fail(PYSA_AMD_ENOTSUP, "block size must be 32 (the MFMA tile) or 0");
  if (a.form == PYSA_AMD_FORM_PLAIN && P->n > pysa::kPlainMaxN)
    return fail(PYSA_AMD_ENOTSUP, "n above the plain kernel's register budget (4096)");
  return 0;
}

template <typename T>
int run_plain(Problem* P, const RunArgs& a, hipStream_t st, double* ms) {
  if (sizeof(T) == 8 && P->n > pysa::kPlainMaxN64)
    return fail(PYSA_AMD_ENOTSUP, "float64 above n = 2048 is not supported by the plain kernel");
  const size_t n = P->n, cnt = (size_t)a.reads * a.R;
  HIP_TRY(P->fields.ensure(cnt * n * sizeof(T)));
  HIP_TRY(P->best_rep_e.ensure(cnt * sizeof(T)));
  HIP_TRY(P->best_rep_sw.ensure(cnt * sizeof(int)));
  if (a.best_sem == 0) HIP_TRY(P->best_rep.ensure(cnt * n));
  pysa::PTParams<T> p;
  p.n = P->n;
  p.ptype = P->ptype;
  p.R = a.R;
  p.S = a.S;
  p.reads = a.reads;
  p.order = a.order;
  p.use_pt = a.use_pt;
  p.logw = a.logw;
  p.best_sem = a.best_sem;
  // f64 field refresh interval: float fields drift with the rounding of
  // millions of row updates; double fields do not (refreshed only at start)
  p.refresh = sizeof(T) == 4 ? 128 : 0;
  p.k0 = (uint32_t)a.seed;
  p.k1 = (uint32_t)(a.seed >> 32);
  p.sweep_offset = a.sweep_offset;
  p.Jt = (const T*)P->dJt;
  p.Jp = (const T*)P->dJp;
  p.ldp = P->ldp;
  const bool sparse = a.form == PYSA_AMD_FORM_SPARSE;
  p.ell_cols = sparse ? P->d_ell_cols : nullptr;
  p.ell_vals = sparse ? (const T*)P->d_ell_vals : nullptr;
  p.ell_d = sparse ? P->ell_d : 0;
  p.h = (const T*)P->dh;
  p.betas = (const T*)a.betas;
  p.bsort = a.bsort;
  p.rep_lr = a.rep_lr;
  p.rep_perm = a.rep_perm;
  p.rep_pt = a.rep_pt;
  p.spins = a.states;
  p.fields = (T*)P->fields.p;
  p.energies = (T*)a.energies;
  p.beta_idx = a.beta_idx;
  p.best_rep = (int8_t*)P->best_rep.p;
  p.best_rep_e = (T*)P->best_rep_e.p;
  p.best_rep_sw = (int*)P->best_rep_sw.p;
  p.best_state = a.best_state;
  p.best_energy = (T*)a.best_energy;
  p.best_sweeps = a.best_sweeps;
  p.log_omegas = a.logw ? a.log_omegas : nullptr;
  p.accepted = (unsigned long long*)a.accepted;
  // Per-read starting values: beta_idx = arange(R) (sa.py:325), the initial
  // state is the first best candidate (simulation.py:143-145).
  const unsigned nb = (unsigned)((cnt + 255) / 256);
  hipLaunchKernelGGL(arange_i64, dim3(nb), dim3(256), 0, st, a.beta_idx, a.reads, a.R);
  hipLaunchKernelGGL((init_best<T>), dim3(nb), dim3(256), 0, st, (const T*)a.energies,
                     (T*)P->best_rep_e.p, (int*)P->best_rep_sw.p, (long long)cnt);
  if (a.best_sem == 0) HIP_TRY(hipMemcpyAsync(P->best_rep.p, a.states, cnt * n, hipMemcpyDeviceToDevice, st));
  if (a.accepted) HIP_TRY(hipMemsetAsync(a.accepted, 0, sizeof(uint64_t) * a.reads, st));
  HIP_TRY(hipGetLastError());
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, st));
  hipError_t le = pysa::launch_plain_pt<T>(p, st);
  HIP_TRY(hipEventRecord(e1, st));
  if (le != hipSuccess) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return fail(PYSA_AMD_EHIP, std::string("plain kernel launch: ") + hipGetErrorString(le));
  }
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
  return 0;
}

template <typename T>
int run_device(Problem* P, const RunArgs& a, hipStream_t st, double* ms) {
  int form = a.form;
  if (form == PYSA_AMD_FORM_AUTO)
    form = pysa::blocked_eligible(P->n, P->dtype, a.R, a.reads, a.order,
                                  a.rep_lr != nullptr || a.rep_pt != nullptr || a.rep_perm != nullptr)
               ? PYSA_AMD_FORM_BLOCKED
               : PYSA_AMD_FORM_PLAIN;
  if (form == PYSA_AMD_FORM_BLOCKED || form == PYSA_AMD_FORM_BLOCKED_EXACT) {
    if (a.rep_perm) return fail(PYSA_AMD_ENOTSUP, "the blocked form does not replay permutations");
    std::string why;
    int rc = pysa::run_blocked(P->blocked, P->dJ, P->dJt, P->dh, P->n, P->ptype, P->dtype, a.reads,
                               a.R, a.S, a.order, a.use_pt, a.logw, a.best_sem,
                               form == PYSA_AMD_FORM_BLOCKED_EXACT ? 3 : 0, a.seed,
                               a.sweep_offset, a.betas, a.bsort, a.rep_lr, a.rep_pt, a.states,
                               a.energies, a.beta_idx, a.best_state, a.best_energy, a.best_sweeps,
                               a.log_omegas, a.accepted, st, ms, &why);
    if (rc != 0) return fail(rc, why);
    return 0;
  }
  if (form == PYSA_AMD_FORM_SPARSE && P->ell_d == 0)
    return fail(PYSA_AMD_ENOTSUP,
                "the sparse form needs a symmetric J with at most n/4 nonzeros per row");
  if (form != PYSA_AMD_FORM_PLAIN && form != PYSA_AMD_FORM_SPARSE)
    return fail(PYSA_AMD_EINVAL, "unknown form");
  RunArgs b = a;
  b.form = form;
  return run_plain<T>(P, b, st, ms);
}

int metropolis_device(int64_t handle, const RunArgs& a, double* kernel_ms, void* stream) {
  Problem* P = get(handle);
  int rc = check_run(P, a);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(P->mu);
  HIP_TRY(hipSetDevice(P->device));
  hipStream_t st = stream ? (hipStream_t)stream : P->stream;
  double ms = 0.0;
  rc = (P->dtype == 0) ? run_device<float>(P, a, st, &ms) : run_device<double>(P, a, st, &ms);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  if (kernel_ms) *kernel_ms = ms;
  return 0;
}

// Device copies of host arrays for the host-pointer entry points.
struct Staged {
  std::vector<void*> ptrs;
  ~Staged() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  hipError_t up(const void* src, size_t bytes, void** dst, hipStream_t st) {
    *dst = nullptr;
    if (!src) return hipSuccess;
    hipError_t e = hipMalloc(dst, bytes ? bytes : 1);
    if (e != hipSuccess) return e;
    ptrs.push_back(*dst);
    return hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, st);
  }
  hipError_t alloc(size_t bytes, void** dst) {
    hipError_t e = hipMalloc(dst, bytes ? bytes : 1);
    if (e == hipSuccess) ptrs.push_back(*dst);
    return e;
  }
};

}  // namespace

namespace {
template <typename T>
__global__ void transpose_kernel(const T* __restrict__ in, T* __restrict__ out, int n) {
  __shared__ T tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int k = ty; k < 32; k += 8)
    if (by + k < n && bx + tx < n) tile[k][tx] = in[(size_t)(by + k) * n + bx + tx];
  __syncthreads();
  for (int k = ty; k < 32; k += 8)
    if (bx + k < n && by + tx < n) out[(size_t)(bx + k) * n + by + tx] = tile[tx][k];
}

// Device copies, transpose and padded rows of a problem whose J and h are
// already in device memory (ordered after earlier work on the default stream).
// The ELL copy of a symmetric host J when every row has at most n/4
// nonzeros (else none: the sparse form is then unavailable).
template <typename T>
hipError_t build_ell(Problem* P, const T* J) {
  const int n = P->n;
  int d = 0;
  for (int i = 0; i < n; ++i) {
    int c = 0;
    for (int j = 0; j < n; ++j) c += J[(size_t)i * n + j] != T(0);
    d = std::max(d, c);
  }
  if (d == 0 || 4 * d > n || d > 64) return hipSuccess;
  std::vector<int> cols((size_t)n * d, -1);
  std::vector<T> vals((size_t)n * d, T(0));
  for (int i = 0; i < n; ++i) {
    int c = 0;
    for (int j = 0; j < n; ++j)
      if (J[(size_t)i * n + j] != T(0)) {
        cols[(size_t)i * d + c] = j;
        vals[(size_t)i * d + c] = J[(size_t)i * n + j];
        ++c;
      }
  }
  hipError_t e = hipMalloc(&P->d_ell_cols, cols.size() * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&P->d_ell_vals, vals.size() * sizeof(T));
  if (e == hipSuccess)
    e = hipMemcpy(P->d_ell_cols, cols.data(), cols.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(P->d_ell_vals, vals.data(), vals.size() * sizeof(T), hipMemcpyHostToDevice);
  if (e == hipSuccess) P->ell_d = d;
  return e;
}

int create_from_device(Problem* P, const void* dJ, const void* dh) {
  const int n = P->n;
  const size_t es = esize(P->dtype), nn = (size_t)n * n;
  hipError_t e = hipMalloc(&P->dJ, nn * es);
  if (e == hipSuccess) e = hipMalloc(&P->dh, (size_t)n * es);
  if (e == hipSuccess) e = hipMalloc(&P->dJt, nn * es);
  if (e == hipSuccess) e = hipMemcpyAsync(P->dJ, dJ, nn * es, hipMemcpyDeviceToDevice, P->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(P->dh, dh, (size_t)n * es, hipMemcpyDeviceToDevice, P->stream);
  if (e == hipSuccess) {
    dim3 grid((n + 31) / 32, (n + 31) / 32);
    if (P->dtype == 0)
      hipLaunchKernelGGL(transpose_kernel<float>, grid, dim3(256), 0, P->stream,
                         (const float*)P->dJ, (float*)P->dJt, n);
    else
      hipLaunchKernelGGL(transpose_kernel<double>, grid, dim3(256), 0, P->stream,
                         (const double*)P->dJ, (double*)P->dJt, n);
    e = hipGetLastError();
  }
  P->symmetric = false;
  if (e == hipSuccess) {
    P->ldp = ((n + 255) / 256) * 256;
    if (P->ldp == n) {
      P->dJp = P->dJt;
    } else {
      const size_t bytes = (size_t)n * P->ldp * es;
      e = hipMalloc(&P->dJp, bytes);
      if (e == hipSuccess) e = hipMemsetAsync(P->dJp, 0, bytes, P->stream);
      if (e == hipSuccess)
        e = hipMemcpy2DAsync(P->dJp, (size_t)P->ldp * es, P->dJt, (size_t)n * es, (size_t)n * es,
                             n, hipMemcpyDeviceToDevice, P->stream);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
  if (e != hipSuccess) return fail(PYSA_AMD_EHIP, std::string("problem copy: ") + hipGetErrorString(e));
  return 0;
}
}  // namespace

extern "C" {

const char* pysa_amd_last_error(void) { return g_err.c_str(); }

const char* pysa_amd_version(void) { return "pysa_amd 0.1.0 (gfx950)"; }


int pysa_amd_problem_create_device(int32_t n, int32_t problem_type, int32_t dtype,
                                   const void* dJ, const void* dh, int64_t* handle_out) {
  g_err.clear();
  if (!handle_out || !dJ || !dh) return fail(PYSA_AMD_EINVAL, "null argument");
  if (n < 1) return fail(PYSA_AMD_EINVAL, "n must be >= 1");
  if (problem_type != 0 && problem_type != 1)
    return fail(PYSA_AMD_EINVAL, "problem_type must be 0 (ising) or 1 (qubo)");
  if (dtype != 0 && dtype != 1) return fail(PYSA_AMD_EINVAL, "dtype must be 0 (f32) or 1 (f64)");
  Problem* P = new Problem();
  P->n = n;
  P->ptype = problem_type;
  P->dtype = dtype;
  hipError_t e = hipGetDevice(&P->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete P;
    return fail(PYSA_AMD_EHIP, std::string("problem stream: ") + hipGetErrorString(e));
  }
  const int rc = create_from_device(P, dJ, dh);
  if (rc) {
    delete P;
    return rc;
  }
  *handle_out = reinterpret_cast<int64_t>(P);
  return 0;
}

int pysa_amd_problem_create(int32_t n, int32_t problem_type, int32_t dtype, const void* J,
                            const void* h, int64_t* handle_out) {
  g_err.clear();
  if (!handle_out || !J || !h) return fail(PYSA_AMD_EINVAL, "null argument");
  if (n < 1) return fail(PYSA_AMD_EINVAL, "n must be >= 1");
  if (problem_type != 0 && problem_type != 1)
    return fail(PYSA_AMD_EINVAL, "problem_type must be 0 (ising) or 1 (qubo)");
  if (dtype != 0 && dtype != 1) return fail(PYSA_AMD_EINVAL, "dtype must be 0 (f32) or 1 (f64)");
  Problem* P = new Problem();
  P->n = n;
  P->ptype = problem_type;
  P->dtype = dtype;
  hipError_t e = hipGetDevice(&P->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking);
  const size_t es = esize(dtype), nn = (size_t)n * n;
  if (e == hipSuccess) e = hipMalloc(&P->dJ, nn * es);
  if (e == hipSuccess) e = hipMalloc(&P->dh, (size_t)n * es);
  if (e == hipSuccess) e = hipMemcpy(P->dJ, J, nn * es, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(P->dh, h, (size_t)n * es, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    bool sym;
    if (dtype == 0) {
      std::vector<float> t;
      sym = transpose_if_needed<float>((const float*)J, n, t);
      if (!sym) {
        e = hipMalloc(&P->dJt, nn * es);
        if (e == hipSuccess) e = hipMemcpy(P->dJt, t.data(), nn * es, hipMemcpyHostToDevice);
      }
    } else {
      std::vector<double> t;
      sym = transpose_if_needed<double>((const double*)J, n, t);
      if (!sym) {
        e = hipMalloc(&P->dJt, nn * es);
        if (e == hipSuccess) e = hipMemcpy(P->dJt, t.data(), nn * es, hipMemcpyHostToDevice);
      }
    }
    P->symmetric = sym;
    if (sym) P->dJt = P->dJ;
    if (sym)
      e = dtype == 0 ? build_ell<float>(P, (const float*)J) : build_ell<double>(P, (const double*)J);
  }
  // Row-padded copy for the plain kernels' 16-byte row loads: whole
  // 256-element chunks per row.
  if (e == hipSuccess) {
    P->ldp = ((n + 255) / 256) * 256;
    if (P->ldp == n) {
      P->dJp = P->dJt;
    } else {
      const size_t bytes = (size_t)n * P->ldp * es;
      e = hipMalloc(&P->dJp, bytes);
      if (e == hipSuccess) e = hipMemset(P->dJp, 0, bytes);
      if (e == hipSuccess)
        e = hipMemcpy2D(P->dJp, (size_t)P->ldp * es, P->dJt, (size_t)n * es, (size_t)n * es, n,
                        hipMemcpyDeviceToDevice);
    }
  }
  if (e != hipSuccess) {
    std::string m = std::string("problem upload: ") + hipGetErrorString(e);
    delete P;
    return fail(PYSA_AMD_EHIP, m);
  }
  *handle_out = reinterpret_cast<int64_t>(P);
  return 0;
}

int pysa_amd_problem_destroy(int64_t handle) {
  g_err.clear();
  Problem* P = get(handle);
  if (!P) return fail(PYSA_AMD_EINVAL, "invalid problem handle");
  delete P;
  return 0;
}

int pysa_amd_metropolis_device(int64_t handle, int32_t reads, int32_t R, int32_t S, int32_t order,
                               int32_t use_pt, int32_t log_omega, int32_t best_semantics,
                               const void* betas, const int32_t* beta_sort, uint64_t seed,
                               uint64_t sweep_offset, int32_t form, int32_t block_b,
                               const double* replay_log_r, const int32_t* replay_perm,
                               const double* replay_pt_logu, int8_t* states, void* energies,
                               int64_t* beta_idx, int8_t* best_state, void* best_energy,
                               int32_t* best_sweeps, double* log_omegas, uint64_t* accepted,
                               double* kernel_ms, void* stream) {
  g_err.clear();
  RunArgs a = {reads, R, S, order, use_pt, log_omega, best_semantics, form, block_b, seed,
               sweep_offset, betas, beta_sort, replay_log_r, replay_perm, replay_pt_logu, states,
               energies, beta_idx, best_state, best_energy, best_sweeps, log_omegas, accepted};
  return metropolis_device(handle, a, kernel_ms, stream);
}

int pysa_amd_metropolis_device_noalloc(int64_t handle, int32_t reads, int32_t R, int32_t S,
                                       int32_t order, int32_t use_pt, int32_t log_omega,
                                       int32_t best_semantics, const void* betas,
                                       const int32_t* beta_sort, uint64_t seed,
                                       uint64_t sweep_offset, int32_t form, int8_t* states,
                                       void* energies, int64_t* beta_idx, int8_t* best_state,
                                       void* best_energy, int32_t* best_sweeps,
                                       double* log_omegas, uint64_t* accepted, void* stream) {
  g_err.clear();
  Problem* P = get(handle);
  RunArgs a = {reads, R, S, order, use_pt, log_omega, best_semantics, form, 0, seed,
               sweep_offset, betas, beta_sort, nullptr, nullptr, nullptr, states,
               energies, beta_idx, best_state, best_energy, best_sweeps, log_omegas, accepted};
  int rc = check_run(P, a);
  if (rc) return rc;
  if (form != PYSA_AMD_FORM_BLOCKED && form != PYSA_AMD_FORM_BLOCKED_EXACT)
    return fail(PYSA_AMD_ENOTSUP, "the allocation-free entry point runs the MFMA form only "
                                  "(PYSA_AMD_FORM_BLOCKED or PYSA_AMD_FORM_BLOCKED_EXACT)");
  std::lock_guard<std::mutex> lk(P->mu);
  HIP_TRY(hipSetDevice(P->device));
  void* sp = nullptr;
  size_t sb = 0;
  std::string why;
  rc = pysa::noalloc_scratch(&sp, &sb, &why);
  if (rc) return fail(rc, why);
  rc = pysa::run_blocked_noalloc(P->blocked, P->dJt, P->dh, P->n, P->ptype, P->dtype, reads, R, S,
                                 order, use_pt, log_omega, best_semantics,
                                 form == PYSA_AMD_FORM_BLOCKED_EXACT ? 3 : 0, seed, sweep_offset,
                                 betas, beta_sort, states, energies, beta_idx, best_state,
                                 best_energy, best_sweeps, log_omegas, accepted, sp, sb,
                                 (hipStream_t)stream, &why);
  if (rc) return fail(rc, why);
  return 0;
}

int pysa_amd_metropolis(int64_t handle, int32_t reads, int32_t R, int32_t S, int32_t order,
                        int32_t use_pt, int32_t log_omega, int32_t best_semantics,
                        const void* betas, const int32_t* beta_sort, uint64_t seed,
                        uint64_t sweep_offset, int32_t form, int32_t block_b,
                        const double* replay_log_r, const int32_t* replay_perm,
                        const double* replay_pt_logu, int8_t* states, void* energies,
                        int64_t* beta_idx, int8_t* best_state, void* best_energy,
                        int32_t* best_sweeps, double* log_omegas, uint64_t* accepted,
                        double* kernel_ms) {
  g_err.clear();
  Problem* P = get(handle);
  RunArgs a = {reads, R, S, order, use_pt, log_omega, best_semantics, form, block_b, seed,
               sweep_offset, betas, beta_sort, replay_log_r, replay_perm, replay_pt_logu, states,
               energies, beta_idx, best_state, best_energy, best_sweeps, log_omegas, accepted};
  int rc = check_run(P, a);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(P->device));
  const size_t es = esize(P->dtype), n = P->n, cnt = (size_t)reads * R;
  const size_t rep = (size_t)reads * S * R * n;
  hipStream_t st = P->stream;
  Staged sg;
  RunArgs d = a;
  void* t;
  HIP_TRY(sg.up(betas, R * es, &t, st));
  d.betas = t;
  HIP_TRY(sg.up(beta_sort, R * sizeof(int32_t), &t, st));
  d.bsort = (const int32_t*)t;
  HIP_TRY(sg.up(replay_log_r, rep * sizeof(double), &t, st));
  d.rep_lr = (const double*)t;
  HIP_TRY(sg.up(replay_perm, rep * sizeof(int32_t), &t, st));
  d.rep_perm = (const int32_t*)t;
  HIP_TRY(sg.up(replay_pt_logu, (size_t)reads * S * (R > 1 ? R - 1 : 1) * sizeof(double), &t, st));
  d.rep_pt = (const double*)t;
  HIP_TRY(sg.up(states, cnt * n, &t, st));
  d.states = (int8_t*)t;
  HIP_TRY(sg.up(energies, cnt * es, &t, st));
  d.energies = t;
  HIP_TRY(sg.alloc(cnt * sizeof(int64_t), &t));
  d.beta_idx = (int64_t*)t;
  HIP_TRY(sg.alloc((size_t)reads * n, &t));
  d.best_state = (int8_t*)t;
  HIP_TRY(sg.alloc((size_t)reads * es, &t));
  d.best_energy = t;
  HIP_TRY(sg.alloc((size_t)reads * sizeof(int32_t), &t));
  d.best_sweeps = (int32_t*)t;
  d.log_omegas = nullptr;
  if (log_omega && log_omegas) {
    HIP_TRY(sg.alloc((size_t)reads * S * sizeof(double), &t));
    d.log_omegas = (double*)t;
  }
  d.accepted = nullptr;
  if (accepted) {
    HIP_TRY(sg.alloc((size_t)reads * sizeof(uint64_t), &t));
    d.accepted = (uint64_t*)t;
  }
  HIP_TRY(hipStreamSynchronize(st));
  rc = metropolis_device(handle, d, kernel_ms, st);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(states, d.states, cnt * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(energies, d.energies, cnt * es, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(beta_idx, d.beta_idx, cnt * sizeof(int64_t), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(best_state, d.best_state, (size_t)reads * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(best_energy, d.best_energy, (size_t)reads * es, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(best_sweeps, d.best_sweeps, (size_t)reads * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (d.log_omegas)
    HIP_TRY(hipMemcpy(log_omegas, d.log_omegas, (size_t)reads * S * sizeof(double), hipMemcpyDeviceToHost));
  if (d.accepted)
    HIP_TRY(hipMemcpy(accepted, d.accepted, (size_t)reads * sizeof(uint64_t), hipMemcpyDeviceToHost));
  return 0;
}

int pysa_amd_ais(int64_t handle, const void* temps, int32_t nt, int32_t chains, int32_t reference,
                 double p, int32_t order, int32_t ordering, uint64_t seed,
                 const double* replay_log_r, const int32_t* replay_perm, int8_t* states,
                 double* log_w, double* energies, int32_t form, double* kernel_ms) {
  g_err.clear();
  Problem* P = get(handle);
  if (!P) return fail(PYSA_AMD_EINVAL, "invalid problem handle");
  if (nt < 1 || chains < 1) return fail(PYSA_AMD_EINVAL, "nt and chains must be >= 1");
  if (!temps || !states || !log_w) return fail(PYSA_AMD_EINVAL, "null array argument");
  if (reference != 0 && reference != 1) return fail(PYSA_AMD_EINVAL, "reference must be 0 or 1");
  if (reference == 1 && !(p > 0.0 && p < 1.0)) return fail(PYSA_AMD_EINVAL, "Bernoulli p must be in (0, 1)");
  if (order < 0 || order > 3) return fail(PYSA_AMD_EINVAL, "order must be 0, 1, 2 or 3");
  if (form != PYSA_AMD_FORM_PLAIN && form != PYSA_AMD_FORM_BLOCKED &&
      form != PYSA_AMD_FORM_BLOCKED_EXACT && form != PYSA_AMD_FORM_AUTO)
    return fail(PYSA_AMD_EINVAL, "AIS form must be PLAIN, BLOCKED, BLOCKED_EXACT or AUTO");
  // The MFMA form: float32, uniform reference, Philox randomness, sequential
  // or block-random order, 65 <= n <= 2048.
  const bool mfma_ok = P->dtype == 0 && reference == 0 && !replay_log_r && !replay_perm &&
                       (order == pysa::kSequential || order == pysa::kBlockRandom) &&
                       P->n > 64 && P->n <= pysa::kBlockedMaxN;
  if ((form == PYSA_AMD_FORM_BLOCKED || form == PYSA_AMD_FORM_BLOCKED_EXACT) && !mfma_ok)
    return fail(PYSA_AMD_ENOTSUP,
                "MFMA AIS needs float32 couplings, 65 <= n <= 2048, sequential or block-random "
                "order, the uniform reference and no replay");
  if (form != PYSA_AMD_FORM_PLAIN && mfma_ok) {
    std::lock_guard<std::mutex> lk(P->mu);
    HIP_TRY(hipSetDevice(P->device));
    hipStream_t st = P->stream;
    const size_t n = P->n;
    Staged sg;
    void *dT, *dS, *dE, *dW;
    HIP_TRY(sg.up(temps, (size_t)nt * sizeof(float), &dT, st));
    HIP_TRY(sg.up(states, (size_t)chains * n, &dS, st));
    HIP_TRY(sg.alloc((size_t)chains * sizeof(double), &dE));
    HIP_TRY(sg.alloc((size_t)chains * sizeof(double), &dW));
    // the initial energies in f64 (simulation.py:301)
    hipError_t le = pysa::launch_energy_f64<float>((const float*)P->dJt, (const float*)P->dh,
                                                   (const int8_t*)dS, (double*)dE, (int)n, chains, st);
    if (le != hipSuccess) {
      (void)hipStreamSynchronize(st);
      return fail(PYSA_AMD_EHIP, std::string("energy launch: ") + hipGetErrorString(le));
    }
    std::string why;
    double ms = 0.0;
    const int rc = pysa::run_blocked_ais(P->blocked, P->dJ, P->dJt, P->dh, (int)n, P->ptype, chains,
                                         nt, (const float*)dT, order, ordering,
                                         form == PYSA_AMD_FORM_BLOCKED_EXACT ? 3 : 0, seed,
                                         (int8_t*)dS, (double*)dE, (double*)dW, st, &ms, &why);
    if (rc != 0) return fail(rc, why);
    if (kernel_ms) *kernel_ms = ms;
    HIP_TRY(hipMemcpy(states, dS, (size_t)chains * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(log_w, dW, (size_t)chains * sizeof(double), hipMemcpyDeviceToHost));
    if (energies)
      HIP_TRY(hipMemcpy(energies, dE, (size_t)chains * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
  if (P->n > pysa::kPlainMaxN) return fail(PYSA_AMD_ENOTSUP, "n above 4096");
  if (P->dtype == 1 && P->n > pysa::kPlainMaxN64)
    return fail(PYSA_AMD_ENOTSUP, "float64 above n = 2048 is not supported by the plain kernel");
  std::lock_guard<std::mutex> lk(P->mu);
  HIP_TRY(hipSetDevice(P->device));
  hipStream_t st = P->stream;
  const size_t es = esize(P->dtype), n = P->n;
  const size_t rep = (size_t)chains * nt * n;
  Staged sg;
  void *dT, *dlr, *dperm, *dS, *dF, *dE, *dW;
  HIP_TRY(sg.up(temps, (size_t)nt * es, &dT, st));
  HIP_TRY(sg.up(replay_log_r, rep * sizeof(double), &dlr, st));
  HIP_TRY(sg.up(replay_perm, rep * sizeof(int32_t), &dperm, st));
  HIP_TRY(sg.up(states, (size_t)chains * n, &dS, st));
  HIP_TRY(sg.alloc((size_t)chains * n * es, &dF));
  HIP_TRY(sg.alloc((size_t)chains * sizeof(double), &dE));
  HIP_TRY(sg.alloc((size_t)chains * sizeof(double), &dW));
  HIP_TRY(hipStreamSynchronize(st));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, st));
  hipError_t le;
  const double lp = reference == 1 ? log(p) : 0.0, l1p = reference == 1 ? log(1.0 - p) : 0.0;
  if (P->dtype == 0) {
    pysa::AISParams<float> a = {(int)n, P->ptype, chains, nt, order, reference, ordering, 256,
                                (uint32_t)seed, (uint32_t)(seed >> 32), lp, l1p,
                                (const float*)P->dJt, (const float*)P->dJp, P->ldp,
                                (const float*)P->dh, (const float*)dT,
                                (const double*)dlr, (const int*)dperm, (int8_t*)dS, (float*)dF,
                                (double*)dE, (double*)dW};
    le = pysa::launch_plain_ais<float>(a, st);
  } else {
    pysa::AISParams<double> a = {(int)n, P->ptype, chains, nt, order, reference, ordering, 0,
                                 (uint32_t)seed, (uint32_t)(seed >> 32), lp, l1p,
                                 (const double*)P->dJt, (const double*)P->dJp, P->ldp,
                                 (const double*)P->dh, (const double*)dT,
                                 (const double*)dlr, (const int*)dperm, (int8_t*)dS, (double*)dF,
                                 (double*)dE, (double*)dW};
    le = pysa::launch_plain_ais<double>(a, st);
  }
  HIP_TRY(hipEventRecord(e1, st));
  if (le != hipSuccess) return fail(PYSA_AMD_EHIP, std::string("AIS launch: ") + hipGetErrorString(le));
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (kernel_ms) *kernel_ms = t;
  HIP_TRY(hipMemcpy(states, dS, (size_t)chains * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(log_w, dW, (size_t)chains * sizeof(double), hipMemcpyDeviceToHost));
  if (energies)
    HIP_TRY(hipMemcpy(energies, dE, (size_t)chains * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int pysa_amd_malloc(uint64_t bytes, void** ptr_out) {
  g_err.clear();
  if (!ptr_out) return fail(PYSA_AMD_EINVAL, "null argument");
  HIP_TRY(hipMalloc(ptr_out, bytes ? bytes : 1));
  return 0;
}

int pysa_amd_free(void* ptr) {
  g_err.clear();
  if (ptr) HIP_TRY(hipFree(ptr));
  return 0;
}

int pysa_amd_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
  g_err.clear();
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

int pysa_amd_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
  g_err.clear();
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int pysa_amd_memcpy_d2d(void* dst, const void* src, uint64_t bytes) {
  g_err.clear();
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return 0;
}

int pysa_amd_device_info(int32_t* compute_units, char* arch, int32_t arch_len) {
  g_err.clear();
  if (!compute_units || (arch_len > 0 && !arch)) return fail(PYSA_AMD_EINVAL, "null argument");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  *compute_units = prop.multiProcessorCount;
  if (arch_len > 0) {
    strncpy(arch, prop.gcnArchName, (size_t)arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  return 0;
}

int pysa_amd_problem_set_shape(int64_t handle, int32_t kernel, int32_t tiles) {
  g_err.clear();
  Problem* P = get(handle);
  if (!P) return fail(PYSA_AMD_EINVAL, "invalid problem handle");
  if (kernel < 0 || kernel > 2)
    return fail(PYSA_AMD_EINVAL, "kernel must be 0 (by size), 1 (pipelined) or 2 (split-K)");
  if (tiles < 0 || tiles > 2) return fail(PYSA_AMD_EINVAL, "tiles must be 0 (by occupancy), 1 or 2");
  std::lock_guard<std::mutex> lk(P->mu);
  P->blocked.shape_kernel = kernel;
  P->blocked.shape_tiles = tiles;
  return 0;
}

int pysa_amd_plain_fields(int64_t handle, int64_t count, void* out) {
  g_err.clear();
  Problem* P = get(handle);
  if (!P) return fail(PYSA_AMD_EINVAL, "invalid problem handle");
  if (count < 0 || (count > 0 && !out)) return fail(PYSA_AMD_EINVAL, "bad arguments");
  std::lock_guard<std::mutex> lk(P->mu);
  const size_t bytes = (size_t)count * P->n * esize(P->dtype);
  if (bytes > P->fields.bytes) return fail(PYSA_AMD_EINVAL, "more fields than the last plain run kept");
  if (count == 0) return 0;
  HIP_TRY(hipSetDevice(P->device));
  HIP_TRY(hipMemcpy(out, P->fields.p, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int pysa_amd_energy(int64_t handle, const int8_t* states, int64_t count, double* out) {
  g_err.clear();
  Problem* P = get(handle);
  if (!P) return fail(PYSA_AMD_EINVAL, "invalid problem handle");
  if (count < 0 || (count > 0 && (!states || !out))) return fail(PYSA_AMD_EINVAL, "bad arguments");
  if (count == 0) return 0;
  std::lock_guard<std::mutex> lk(P->mu);
  HIP_TRY(hipSetDevice(P->device));
  hipStream_t st = P->stream;
  Staged sg;
  void *dS, *dO;
  HIP_TRY(sg.up(states, (size_t)count * P->n, &dS, st));
  HIP_TRY(sg.alloc((size_t)count * sizeof(double), &dO));
  hipError_t le = (P->dtype == 0)
                      ? pysa::launch_energy_f64<float>((const float*)P->dJt, (const float*)P->dh,
                                                       (const int8_t*)dS, (double*)dO, P->n, count, st)
                      : pysa::launch_energy_f64<double>((const double*)P->dJt, (const double*)P->dh,
                                                        (const int8_t*)dS, (double*)dO, P->n, count, st);
  if (le != hipSuccess) return fail(PYSA_AMD_EHIP, std::string("energy launch: ") + hipGetErrorString(le));
  HIP_TRY(hipMemcpyAsync(out, dO, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return 0;
}

}  // extern "C"

extern "C" int pysa_amd_batched_device(int32_t count, const int32_t* n, const int32_t* ld,
                                       const int64_t* j_off, const int8_t* signs,
                                       const float* scale, const float* h, const int64_t* h_off,
                                       int32_t problem_type, int32_t reads, int32_t R, int32_t S,
                                       int32_t order, int32_t use_pt, int32_t best_semantics,
                                       const float* betas, uint64_t seed, uint64_t sweep_offset,
                                       const int64_t* s_off, int8_t* states, float* energies,
                                       int64_t* beta_idx, int8_t* best_state, float* best_energy,
                                       int32_t* best_sweeps, uint64_t* accepted, int32_t form,
                                       double* kernel_ms, void* stream) {
  g_err.clear();
  if (count < 1 || reads < 1 || R < 1 || R > 16 || S < 1 || order < 0 || order > 3 ||
      problem_type < 0 || problem_type > 1 || best_semantics < 0 || best_semantics > 1)
    return fail(PYSA_AMD_EINVAL, "batched: count, reads >= 1, 1 <= R <= 16, S >= 1, order 0-3");
  if (form != PYSA_AMD_FORM_PLAIN && form != PYSA_AMD_FORM_BLOCKED && form != PYSA_AMD_FORM_AUTO)
    return fail(PYSA_AMD_EINVAL, "batched: form must be PLAIN, BLOCKED or AUTO");
  const bool mfma_order = order == pysa::kSequential || order == pysa::kBlockRandom;
  if (form == PYSA_AMD_FORM_BLOCKED && !mfma_order)
    return fail(PYSA_AMD_ENOTSUP, "batched MFMA form: sequential or block-random order only");
  if (!n || !ld || !j_off || !signs || !scale || !betas || !s_off || !states || !energies ||
      !beta_idx || !best_state || !best_energy || !best_sweeps || (h && !h_off))
    return fail(PYSA_AMD_EINVAL, "batched: null array");
  if ((int64_t)count * reads > 0x7fffffffLL) return fail(PYSA_AMD_EINVAL, "batched: too many reads");
  hipStream_t st = (hipStream_t)stream;
  std::vector<int32_t> hn(count), hld(count);
  std::vector<int64_t> hj(count), hs(count), hho(h ? count : 0);
  std::vector<float> hsc(count);
  HIP_TRY(hipMemcpyAsync(hn.data(), n, sizeof(int32_t) * count, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hld.data(), ld, sizeof(int32_t) * count, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hj.data(), j_off, sizeof(int64_t) * count, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hs.data(), s_off, sizeof(int64_t) * count, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hsc.data(), scale, sizeof(float) * count, hipMemcpyDeviceToHost, st));
  if (h) HIP_TRY(hipMemcpyAsync(hho.data(), h_off, sizeof(int64_t) * count, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  int max_n = 0;
  int64_t acc_states = 0;
  for (int k = 0; k < count; ++k) {
    if (hn[k] < 1 || hn[k] > pysa::kBatchedMaxN)
      return fail(PYSA_AMD_EINVAL, "batched: instance size out of [1, 2048]");
    if (hld[k] < hn[k] || hld[k] % 16 != 0 || hj[k] < 0 || hj[k] % 16 != 0)
      return fail(PYSA_AMD_EINVAL, "batched: ld must be >= n and a multiple of 16, j_off 16-aligned");
    if (hs[k] != acc_states) return fail(PYSA_AMD_EINVAL, "batched: s_off must be the prefix sum of reads*R*n");
    acc_states += (int64_t)reads * R * hn[k];
    max_n = hn[k] > max_n ? hn[k] : max_n;
  }
  if (max_n > 1024 && R > 8) return fail(PYSA_AMD_EINVAL, "batched: R <= 8 when an instance has n > 1024");
  // Largest instances first, so the short ones fill the tail of the launch.
  // The MFMA form (BLOCKED / AUTO, sequential or block-random order) takes
  // the instances of three blocks or more; the plain kernel the others.
  std::vector<int> idx(count), mf, pl;
  for (int k = 0; k < count; ++k) idx[k] = k;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return hn[a] > hn[b]; });
  const bool use_mfma = form != PYSA_AMD_FORM_PLAIN && mfma_order;
  for (int k : idx) (use_mfma && hn[k] >= pysa::kBatchedMfmaMinN ? mf : pl).push_back(k);
  int max_n_pl = 0;
  for (int k : pl) max_n_pl = hn[k] > max_n_pl ? hn[k] : max_n_pl;
  const int items = (int)pl.size() * reads;
  std::vector<int32_t> om(items > 0 ? items : 1);
  for (int q = 0, o = 0; q < (int)pl.size(); ++q)
    for (int r = 0; r < reads; ++r) om[o++] = pl[q] * reads + r;
  void *d_om = nullptr, *d_best = nullptr;
  HIP_TRY(hipMalloc(&d_om, sizeof(int32_t) * om.size()));
  if (best_semantics == 0 && hipMalloc(&d_best, (size_t)acc_states) != hipSuccess) {
    (void)hipFree(d_om);
    return fail(PYSA_AMD_EHIP, "batched: allocating best-state scratch");
  }
  struct Frees {
    void* a;
    void* b;
    ~Frees() {
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
    }
  } frees{d_om, d_best};
  HIP_TRY(hipMemcpyAsync(d_om, om.data(), sizeof(int32_t) * om.size(), hipMemcpyHostToDevice, st));
  if (accepted) HIP_TRY(hipMemsetAsync(accepted, 0, sizeof(uint64_t) * count * reads, st));
  // the MFMA kernel's per-replica bests start from the initial states
  if (best_semantics == 0 && !mf.empty())
    HIP_TRY(hipMemcpyAsync(d_best, states, (size_t)acc_states, hipMemcpyDeviceToDevice, st));
  pysa::BatchParams p;
  p.count = count;
  p.reads = reads;
  p.R = R;
  p.S = S;
  p.order = order;
  p.use_pt = use_pt ? 1 : 0;
  p.best_sem = best_semantics;
  p.ptype = problem_type;
  p.k0 = (uint32_t)seed;
  p.k1 = (uint32_t)(seed >> 32);
  p.sweep_offset = sweep_offset;
  p.inst_n = n;
  p.ld = ld;
  p.j_off = j_off;
  p.signs = signs;
  p.scale = scale;
  p.h = h;
  p.h_off = h_off;
  p.betas = betas;
  p.s_off = s_off;
  p.order_map = (const int32_t*)d_om;
  p.states = states;
  p.energies = energies;
  p.beta_idx = beta_idx;
  p.best_state = best_state;
  p.best_energy = best_energy;
  p.best_sweeps = best_sweeps;
  p.accepted = (unsigned long long*)accepted;
  p.best_rep = (int8_t*)d_best;
  pysa::BatchedMfmaRun mrun;  // its buffers are freed after the synchronisation below
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(PYSA_AMD_EHIP, "batched: creating events");
  }
  (void)hipEventRecord(e0, st);
  hipError_t le = mrun.enqueue(p, mf, hn.data(), hld.data(), hj.data(), hs.data(),
                               h ? hho.data() : nullptr, hsc.data(), st);
  if (le == hipSuccess && items > 0) le = pysa::launch_batched(p, items, max_n_pl, st);
  (void)hipEventRecord(e1, st);
  if (le != hipSuccess) (void)hipStreamSynchronize(st);  // nothing in flight when buffers are freed
  hipError_t se = (le == hipSuccess) ? hipEventSynchronize(e1) : le;
  float t = 0.f;
  if (se == hipSuccess) (void)hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (le != hipSuccess) return fail(PYSA_AMD_EHIP, std::string("batched launch: ") + hipGetErrorString(le));
  if (se != hipSuccess) return fail(PYSA_AMD_EHIP, std::string("batched run: ") + hipGetErrorString(se));
  if (kernel_ms) *kernel_ms = t;
  return 0;
}
