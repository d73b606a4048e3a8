/*
 * pysa_amd -- C ABI of the MI355X Metropolis / parallel-tempering engine.
 *
 * The reference (nasa/PySA) is pure Python + numba: its hot path is bound by
 * passing numba functions by value, not through an FFI.  These entry points
 * replace that path below the public Python API:
 *
 *   Solver.__init__ split/cast          pysa/sa.py:37-70, 204-205
 *   simulation_parallel / _sequential   pysa/simulation.py:123-236
 *     + random_sweep / sequential_sweep pysa/simulation.py:41-76
 *     + ising/qubo update_spin          pysa/ising.py:40-63, pysa/qubo.py:40-63
 *     + pt                              pysa/utils.py:26-58
 *     + get_log_omega                   pysa/ais.py:79-112
 *   per-read loop                       pysa/sa.py:425-427
 *   ais_simulation_parallel/ais_sample  pysa/simulation.py:266-317
 *   ising/qubo get_energy               pysa/ising.py:32-37, pysa/qubo.py:32-37
 *
 * Conventions
 *   - Every function returns 0 on success and a negative PYSA_AMD_E* code on
 *     failure; pysa_amd_last_error() then holds a message (thread-local).
 *   - Arrays are C-contiguous.  Host-pointer entry points copy in and out and
 *     keep no pointer after they return.  *_device entry points take device
 *     pointers and run on the given HIP stream (NULL = the problem's stream);
 *     they return after the work has completed.
 *   - dtype 0 = float32, 1 = float64 ("float_type" of the reference).
 *     Energy and beta arrays use the problem's dtype; spin arrays are int8
 *     holding +1/-1 (ising) or 1/0 (qubo).
 *   - problem_type 0 = ising, 1 = qubo.
 *   - order 0 = sequential, 1 = random (a fresh permutation per chain and
 *     sweep), 2 = random_shared (one permutation per sweep shared by every
 *     chain, the order the blocked form uses).
 */
#ifndef PYSA_AMD_H
#define PYSA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYSA_AMD_OK 0
#define PYSA_AMD_EINVAL (-1)  /* bad argument */
#define PYSA_AMD_EHIP (-2)    /* HIP runtime error */
#define PYSA_AMD_ENOTSUP (-3) /* valid request this build does not support */

/* Engine forms accepted by pysa_amd_metropolis*.  BLOCKED encodes real
 * couplings as hi+lo bf16 pieces (integer couplings exactly); BLOCKED_EXACT
 * uses hi+mid+lo pieces, exact for every float32. */
#define PYSA_AMD_FORM_PLAIN 0
#define PYSA_AMD_FORM_BLOCKED 1
#define PYSA_AMD_FORM_AUTO 2
#define PYSA_AMD_FORM_BLOCKED_EXACT 3
#define PYSA_AMD_FORM_SPARSE 4

/* Message describing the last failure on this thread ("" if none). */
const char* pysa_amd_last_error(void);

/* Library version string. */
const char* pysa_amd_version(void);

/*
 * Upload one dense problem and keep it resident on the current device.
 * J: n*n couplings with zero diagonal (row i holds J[i][:]); h: n local
 * fields.  Replaces the split and cast of Solver.__init__ / metropolis_update
 * (pysa/sa.py:67-70, 204-205): the caller passes the already-split arrays.
 */
int pysa_amd_problem_create(int32_t n, int32_t problem_type, int32_t dtype, const void* J,
                            const void* h, int64_t* handle_out);

/*
 * The same from device memory (J and h already resident, e.g. torch tensors
 * of a training loop): no host copy.  J need not be symmetric (its transpose
 * is formed on the device).  The caller makes sure the work that wrote J and
 * h has completed (synchronise its stream); the call returns when the
 * problem's copies are made.  Serves the torch-native negative phase of RBM
 * training (tutorials/rbm_pysa.py:111-140 rebuilds a Solver from host arrays).
 */
int pysa_amd_problem_create_device(int32_t n, int32_t problem_type, int32_t dtype,
                                   const void* dJ, const void* dh, int64_t* handle_out);

/* Release a problem and its device memory. */
int pysa_amd_problem_destroy(int64_t handle);

/*
 * Metropolis sweeps with optional parallel tempering for `reads` independent
 * reads of R replicas (pysa/simulation.py:123-236 called once per read by
 * pysa/sa.py:425-427).
 *
 *   S                sweeps per read (>= 1)
 *   use_pt           replica exchange after every sweep (pysa/utils.py:26-58)
 *   log_omega        compute the per-sweep log-omega (pysa/ais.py:79-112);
 *                    requires a zero beta, validated by the caller
 *   best_semantics   0 = per-replica best then argmin (simulation_parallel),
 *                    1 = global best in replica order (simulation_sequential)
 *   betas            R inverse temperatures in slot order (dtype)
 *   beta_sort        R slot indices sorted by ascending beta (log_omega only)
 *   seed             Philox key; sweep_offset shifts the sweep counter so a
 *                    resumed run continues the same stream
 *   form             PYSA_AMD_FORM_*; AUTO picks BLOCKED where it applies
 *                    (float32, sequential or block-random order, no replay,
 *                    R <= 64, 32 <= n <= 2048), else PLAIN.  SPARSE is the
 *                    plain form with an accepted flip updating only the
 *                    nonzeros of its coupling row (same results bit for
 *                    bit); it needs a symmetric J with at most min(n/4, 64)
 *                    nonzeros per row, whose ELL copy pysa_amd_problem_create
 *                    builds
 *   block_b          blocked-form block size: 0 or 32 (the MFMA tile; other
 *                    values give PYSA_AMD_ENOTSUP)
 *   replay_log_r     NULL, or reads*S*R*n log-uniforms indexed
 *                    [read][sweep][replica][position] (replay mode)
 *   replay_perm      NULL, or reads*S*R*n visit orders [read][sweep][replica][i]
 *   replay_pt_logu   NULL, or reads*S*(R-1) PT log-uniforms [read][sweep][pair],
 *                    pair 0 being (R-1, R-2)
 *   states           in/out reads*R*n spins;  energies in/out reads*R (dtype)
 *   beta_idx         out reads*R: slot -> replica map after the last sweep
 *   best_state       out reads*n; best_energy out reads (dtype);
 *   best_sweeps      out reads (0-based sweep of the best state)
 *   log_omegas       out reads*S (f64) or NULL
 *   accepted         out reads: accepted flips, or NULL
 *   kernel_ms        out: device time of the run (HIP events), or NULL
 */
int pysa_amd_metropolis(int64_t handle, int32_t reads, int32_t R, int32_t S, int32_t order,
                        int32_t use_pt, int32_t log_omega, int32_t best_semantics,
                        const void* betas, const int32_t* beta_sort, uint64_t seed,
                        uint64_t sweep_offset, int32_t form, int32_t block_b,
                        const double* replay_log_r, const int32_t* replay_perm,
                        const double* replay_pt_logu, int8_t* states, void* energies,
                        int64_t* beta_idx, int8_t* best_state, void* best_energy,
                        int32_t* best_sweeps, double* log_omegas, uint64_t* accepted,
                        double* kernel_ms);

/* Same as pysa_amd_metropolis with every array (including betas, beta_sort
 * and the replay arrays) already in device memory. */
int pysa_amd_metropolis_device(int64_t handle, int32_t reads, int32_t R, int32_t S, int32_t order,
                               int32_t use_pt, int32_t log_omega, int32_t best_semantics,
                               const void* betas, const int32_t* beta_sort, uint64_t seed,
                               uint64_t sweep_offset, int32_t form, int32_t block_b,
                               const double* replay_log_r, const int32_t* replay_perm,
                               const double* replay_pt_logu, int8_t* states, void* energies,
                               int64_t* beta_idx, int8_t* best_state, void* best_energy,
                               int32_t* best_sweeps, double* log_omegas, uint64_t* accepted,
                               double* kernel_ms, void* stream);

/*
 * pysa_amd_metropolis_device for the MFMA form (PYSA_AMD_FORM_BLOCKED or
 * _BLOCKED_EXACT) that allocates nothing: no device memory, no event, no
 * synchronisation.  No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).
 *   - The problem's couplings must already be encoded for the form: run the
 *     same form once through pysa_amd_metropolis(_device) on the handle.
 *   - The per-replica best states live in a static device array of 32 MiB
 *     per GPU (held until the process ends): reads * R * (n + 8) bytes must
 *     fit, else PYSA_AMD_EINVAL and nothing is launched.
 *   - The work is enqueued on `stream` (NULL = the default stream) and the
 *     call returns at once; the caller synchronises.  Calls that overlap on
 *     one GPU share that array and race.
 * Arguments as in pysa_amd_metropolis_device (no replay, no timing).
 */
int pysa_amd_metropolis_device_noalloc(int64_t handle, int32_t reads, int32_t R, int32_t S,
                                       int32_t order, int32_t use_pt, int32_t log_omega,
                                       int32_t best_semantics, const void* betas,
                                       const int32_t* beta_sort, uint64_t seed,
                                       uint64_t sweep_offset, int32_t form, int8_t* states,
                                       void* energies, int64_t* beta_idx, int8_t* best_state,
                                       void* best_energy, int32_t* best_sweeps,
                                       double* log_omegas, uint64_t* accepted, void* stream);

/*
 * Annealed importance sampling over `chains` independent chains
 * (pysa/simulation.py:266-317 with update_spin_ref, pysa/ising.py:66-92).
 * The caller passes J and h already scaled by the target beta
 * (pysa/sa.py:511-512) in the problem handle.
 *
 *   temps        nt inverse temperatures (dtype), (t+1)/nt in the reference
 *   reference    0 = uniform (pysa/ais.py:137-169), 1 = Bernoulli with
 *                P(spin = -1 or 0) = p (pysa/ais.py:175-226)
 *   ordering     0 = weight taken after the sweep at beta_t (reference,
 *                pysa/simulation.py:305-316), 1 = before it (Neal 2001)
 *   states       in/out chains*n spins
 *   log_w        out chains log-weights (f64, without log Z0)
 *   energies     out chains final energies as tracked (f64: recomputed with
 *                the fields every 256 steps (plain) or 64 steps (MFMA form) for
 *                float32 couplings, updated by the accepted changes in
 *                between), or NULL
 *   form         PYSA_AMD_FORM_PLAIN: the plain AIS kernel (every case);
 *                BLOCKED / BLOCKED_EXACT: the MFMA form (2 or 3 bf16 pieces,
 *                int8 for small integer couplings) for float32, the uniform
 *                reference, no replay, order 0 or 3 and 65 <= n <= 2048
 *                (PYSA_AMD_ENOTSUP otherwise); AUTO: BLOCKED where it applies,
 *                else PLAIN.  With exactly encoded couplings every form gives
 *                the same states and weights for the same order.
 */
int pysa_amd_ais(int64_t handle, const void* temps, int32_t nt, int32_t chains, int32_t reference,
                 double p, int32_t order, int32_t ordering, uint64_t seed,
                 const double* replay_log_r, const int32_t* replay_perm, int8_t* states,
                 double* log_w, double* energies, int32_t form, double* kernel_ms);

/*
 * Device-memory helpers for callers without their own HIP allocator (the
 * benchmark keeps its inputs resident in HBM with these).  No counterpart in
 * the reference, which has no device memory.
 */
int pysa_amd_malloc(uint64_t bytes, void** ptr_out);
int pysa_amd_free(void* ptr);
int pysa_amd_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
int pysa_amd_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
int pysa_amd_memcpy_d2d(void* dst, const void* src, uint64_t bytes);

/* E = s.(J s / 2 + h) in f64 for `count` states (pysa/ising.py:32-37). */
int pysa_amd_energy(int64_t handle, const int8_t* states, int64_t count, double* out);

/*
 * The plain form's resident local fields F = J s (without h) after the last
 * plain-form run on this problem: the first `count` chains (read-major,
 * count*n values in the problem dtype).  No counterpart in the reference,
 * which recomputes every field from the coupling row; it measures the drift
 * that the periodic f64 recomputation bounds.
 */
int pysa_amd_plain_fields(int64_t handle, int64_t count, void* out);

/*
 * Launch shape of the MFMA form on this problem, for the autotuner
 * (pysa_amd/tune.py).  No counterpart in the reference, whose numba loop has
 * no launch shape.  Results do not depend on the shape, only the speed.
 *   kernel  0 = by size (the pipelined kernel from three 32-spin blocks up),
 *           1 = pipelined, 2 = split-K (the block-random order always runs
 *           the pipelined kernel)
 *   tiles   replicas per workgroup of the pipelined kernel in tiles of 32:
 *           0 = by occupancy (one tile when two-tile workgroups would leave
 *           compute units idle), 1 or 2 (R > 32 always takes 2)
 */
int pysa_amd_problem_set_shape(int64_t handle, int32_t kernel, int32_t tiles);

/*
 * The current device's compute units and architecture name (arch_len bytes,
 * NUL-terminated), the device part of the autotuner's cache key.  No
 * counterpart in the reference.
 */
int pysa_amd_device_info(int32_t* compute_units, char* arch, int32_t arch_len);

/*
 * Many independent instances of different sizes in one launch.  PySA has no
 * batched entry: a batch of problems is a caller's Python loop over Solver
 * objects, each running pysa/sa.py:92-430 with its per-read loop
 * (pysa/sa.py:425-427) over pysa/simulation.py:123-236.  This replaces that
 * whole loop for float32 problems whose couplings are an integer matrix times
 * a per-instance scale (e.g. Sherrington-Kirkpatrick +-1/sqrt(n)).
 *
 * Instance k (0 <= k < count), all arrays in device memory:
 *   n[k]        spins, 1..2048
 *   signs + j_off[k]  int8 S_k, n[k] rows of ld[k] bytes (ld a multiple of 16,
 *               j_off 16-aligned, padding zero); J_k = scale[k] * S_k must be
 *               symmetric with a zero diagonal
 *   h + h_off[k]  n[k] float32 local fields, or h == NULL for h = 0
 *   betas + k*R   its ladder (slot order, float32)
 *   s_off[k]    sum_{j<k} reads*R*n[j]: offset of its states
 *               (reads*R*n[k] int8 at states + s_off[k]) and, divided by R,
 *               of its best states (reads*n[k] int8 at best_state + s_off[k]/R)
 * Per (instance, read) g = k*reads + read: energies/beta_idx [g*R ..],
 * best_energy/best_sweeps/accepted [g].  Local fields are exact integers
 * S_k s; the field of spin i is the float32 value scale*c_i + h_i (the
 * product rounded, then the sum).  Randomness is the Philox stream of
 * pysa_amd_metropolis with read index g.
 *   R <= 16 (R <= 8 when some n > 1024), order 0-3, best_semantics as in
 *   pysa_amd_metropolis.
 *   form  PYSA_AMD_FORM_PLAIN: one workgroup per (instance, read), the fields
 *         in registers; BLOCKED or AUTO: instances with n >= 65 run the MFMA
 *         form (int8 products of 32-spin blocks, one workgroup per instance
 *         and up to 32 or 64 chains), the others the plain kernel.  BLOCKED
 *         needs order 0 or 3 (PYSA_AMD_ENOTSUP otherwise); AUTO falls back
 *         to PLAIN for orders 1-2.  Every form gives the same results for
 *         the same order.
 * kernel_ms: device time of the call's work (packing included).
 */
int pysa_amd_batched_device(int32_t count, const int32_t* n, const int32_t* ld,
                            const int64_t* j_off, const int8_t* signs, const float* scale,
                            const float* h, const int64_t* h_off, int32_t problem_type,
                            int32_t reads, int32_t R, int32_t S, int32_t order, int32_t use_pt,
                            int32_t best_semantics, const float* betas, uint64_t seed,
                            uint64_t sweep_offset, const int64_t* s_off, int8_t* states,
                            float* energies, int64_t* beta_idx, int8_t* best_state,
                            float* best_energy, int32_t* best_sweeps, uint64_t* accepted,
                            int32_t form, double* kernel_ms, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYSA_AMD_H */
