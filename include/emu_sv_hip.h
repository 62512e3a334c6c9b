/*
 * emu_sv_hip.h -- C ABI of the MI355X (gfx950) state-vector Krylov engine.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and never
 * allocates or frees a caller buffer.  Device pointers are owned by the
 * caller (e.g. torch tensors).  Calls that take a hipStream_t only enqueue
 * work on that stream; the caller synchronises.  Return value: 0 on success,
 * a negative EMU_SV_E* code otherwise (emu_sv_last_error() gives the text).
 *
 * Conventions shared with the reference emulator (paths relative to the
 * reference checkout):
 *   - state: 1-D contiguous complex vector of length D = 2^N, interleaved
 *     (re, im); complex128 ("c128") or complex64 ("c64"); every device
 *     vector is 16-byte aligned (kernels move 16-byte packets).
 *   - qubit q <-> bit (N-1-q) of the amplitude index: qubit 0 is the most
 *     significant bit (emu_sv/hamiltonian.py:107); bit value 1 = Rydberg "r".
 *   - H = sum_q (O_q/2)[cos(phi_q) X_q + sin(phi_q) Y_q] - sum_q d_q n_q
 *         + sum_{i<j} U_ij n_i n_j   (emu_sv/hamiltonian.py:11); only the
 *     upper triangle U[i][j], i<j, of the interaction matrix is read
 *     (emu_sv/hamiltonian.py:157-161).
 *   - the Krylov operator is op(x) = sigma * H x with sigma = -i dt
 *     (emu_sv/time_evolution/torch.py:161-162).
 */
#ifndef EMU_SV_HIP_H
#define EMU_SV_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define EMU_SV_OK 0
#define EMU_SV_EINVAL -1        /* bad size / shape / argument -> ValueError   */
#define EMU_SV_EHIP -2          /* HIP runtime error           -> RuntimeError */
#define EMU_SV_ENOTCONVERGED -3 /* Krylov not converged        -> RecursionError */

/* element types */
#define EMU_SV_C128 0
#define EMU_SV_C64 1

/* emu_sv_hpsi flags: bit 0 complex coefficients, bit 1 noise, bit 2 keep
 * the diagonal table, bits 8-11 kernel kind */
#define EMU_SV_FLAG_COMPLEX 1 /* some phi != 0: complex flip coefficients */
#define EMU_SV_FLAG_NOISE 2   /* the block carries the non-Hermitian term   */
                              /* -i/2 sum L^+L (emu_sv_pack_noise); implies */
                              /* complex coefficients; the LDS kernel, or   */
                              /* the btile kernel for bare H.psi above one  */
                              /* big tile (other kinds run the LDS kernel)  */
/* The first pass's diagonal constants are tabulated in the workspace by a
 * small kernel before the passes.  With EMU_SV_FLAG_KEEP_DIAG (emu_sv_hpsi,
 * emu_sv_lanczos_matvec, emu_sv_lanczos_matvec_dev) the table is reused when
 * it was built from a block of the same pack generation (block slot 1135),
 * n and first-pass tile size; otherwise it is rebuilt.  The check runs on
 * the device (one stamp per table row), so a freed block whose address is
 * reused by another block never reuses a stale table.  A block modified in
 * place after packing must be repacked (new generation) to be seen as new.
 * The Krylov loop of one time step sets it on every iteration after the
 * first. */
#define EMU_SV_FLAG_KEEP_DIAG 4
#define EMU_SV_KERNEL_DEFAULT (0 << 8) /* library choice (the wave kernel)  */
#define EMU_SV_KERNEL_LDS (1 << 8)     /* workgroup tiles staged in LDS     */
#define EMU_SV_KERNEL_WAVE (2 << 8)    /* wavefront tiles in registers,     */
                                       /* cross-lane flips; tile_log 9..11  */
#define EMU_SV_KERNEL_TILE (3 << 8)    /* contiguous first pass + one pass  */
                                       /* of c contiguous low bits (min_contig) */
                                       /* and the next high bits; tile_log  */
                                       /* 11..13 (c128), 11..14 (c64)        */
#define EMU_SV_KERNEL_BTILE (4 << 8)   /* big register tiles (1024 threads, */
                                       /* 2^14 c128 / 2^15 c64 amplitudes): */
                                       /* a contiguous first pass, then     */
                                       /* passes of min_contig low bits and */
                                       /* the next high bits (tile_log is   */
                                       /* ignored); two passes up to n = 26 */
/* Tile kernel only: the last pass reads the flip partners of up to P
 * leftover top bits directly (orbit-ordered tiles; 2, 4 or 6 of them are
 * compiled), P in flags bits 12-15; the first (contiguous) pass may use its
 * own tile log t (flags bits 16-19, 0 = tile_log). */
#define EMU_SV_TILE_PARTNERS(p) (((p) & 0xF) << 12)
#define EMU_SV_TILE_FIRST_LOG(t) (((t) & 0xF) << 16)

#define EMU_SV_MAX_QUBITS 32
/* Packed Hamiltonian parameter block, indexed by BIT position b = N-1-q:
 *   [0,32)        Re c_b, c_b = (O_q/2) e^{i phi_q}  (<1|H|0> on that qubit)
 *   [32,64)       Im c_b
 *   [64,96)       d_b (detuning)
 *   [96,96+1024)  symmetric U_b[b][b'] (upper triangle of the qubit matrix
 *                 mirrored, zero diagonal)
 *   [1120]        flag: 1.0 if any Im c_b != 0 (the reference's complex
 *                 sigma path, emu_sv/hamiltonian.py:61)
 *   [1121]        flag: 1.0 if the noise term below is present
 *   [1122,1130)   noise 2x2 complex K = -i/2 sum_l L_l^+ L_l, row-major
 *                 (re, im) pairs K00, K01, K10, K11 in the (g, r) basis;
 *                 applied to every qubit: <1|H|0> += K10, <0|H|1> += K01,
 *                 diagonal += K00 per g bit and K11 per r bit
 *                 (emu_sv/hamiltonian.py:61-67, 147-151)
 *   [1130,1135)   reserved (zero)
 *   [1135]        pack generation: a process-unique positive integer
 *                 written by emu_sv_pack_hamiltonian.  EMU_SV_FLAG_KEEP_DIAG
 *                 reuses a workspace's first-pass diagonal table only when
 *                 the table was built from a block of the same generation,
 *                 n and tile (checked on the device); a block of generation
 *                 0 (not packed by the library) is always rebuilt.      */
#define EMU_SV_HAM_DOUBLES 1136

/* Last error message of the calling thread ("" if none). */
const char *emu_sv_last_error(void);

/* Library version string. */
const char *emu_sv_version(void);

/* Pack (O, d, phi, U) given per QUBIT into the parameter block above
 * (host memory).  omega/delta/phi: n doubles; interaction: n*n row-major
 * doubles (only i<j read).  Replaces RydbergHamiltonian.__init__ and the
 * diagonal it stores (emu_sv/hamiltonian.py:42-67, 139-162): the diagonal is
 * never materialised, kernels rebuild it on the fly from this block. */
int emu_sv_pack_hamiltonian(int n, const double *omega, const double *delta,
                            const double *phi, const double *interaction,
                            double *block_out);

/* Add the single-qubit noise term K = -i/2 sum_l L_l^+ L_l (2x2 complex,
 * 8 doubles row-major (re, im), host) to a packed block (host) and set its
 * noise flag.  Replaces the noise part of RydbergHamiltonian.__init__ and
 * _create_diagonal (emu_sv/hamiltonian.py:61-67, 147-151); K itself is
 * compute_noise_from_lindbladians (emu_base/jump_lindblad_operators.py:80-93).
 * Kernels given EMU_SV_FLAG_NOISE apply it on the fly. */
int emu_sv_pack_noise(const double *noise, double *block_inout);

/* Number of H.psi passes the planner uses for n qubits, element type and
 * schedule (each pass streams the vector once). */
int emu_sv_plan_passes(int n, int dtype, int tile_log, int min_contig, int flags);

/* Bytes of the per-workgroup partial-sum workspace needed by the calls
 * below for a vector of 2^n elements. */
size_t emu_sv_workspace_bytes(int n, int dtype);

/* w = sigma*H*(v_scale*v)  [+ reductions].  Replaces
 * RydbergHamiltonian.__mul__ (emu_sv/hamiltonian.py:69-137) composed with the
 * -1j*dt scaling of op (time_evolution/torch.py:161-162) and fused with the
 * dot products of one Lanczos iteration (emu_base/math/krylov_exp.py:52-58).
 *   ham      device pointer to a packed parameter block
 *   v, w     device vectors (must not alias); prev may be NULL
 *   dots_out device pointer to 7 doubles receiving, after the call:
 *            [0] ||w||^2, [1,2] <v|w>, [3,4] <prev|w>, [5,6] <v|prev>
 *            (v and prev taken with their scales); NULL = no reductions.
 *   workspace device scratch of emu_sv_workspace_bytes(n, dtype) bytes,
 *            zero-filled before its first use (it holds a counter that the
 *            library re-arms after every reduction, and the first pass's
 *            diagonal table); required, also without reductions.
 *   tile_log / min_contig: log2 of the tile (the amplitudes one workgroup
 *            (LDS kernel) or one wavefront (wave kernel) owns) and the
 *            minimum contiguous low bits of the high-bit passes; 0 = default.
 *   flags    EMU_SV_FLAG_COMPLEX when any phi != 0 (emu_sv/hamiltonian.py:61),
 *            or'ed with an EMU_SV_KERNEL_* kind.
 */
int emu_sv_hpsi(int n, int dtype, const double *ham, double sigma_re,
                double sigma_im, const void *v, double v_scale,
                const void *prev, double prev_scale, void *w, double *dots_out,
                void *workspace, int tile_log, int min_contig, int flags,
                hipStream_t stream);

/* emu_sv_hpsi without a workspace argument: it allocates nothing and uses
 * instead a static device array of the library (24 MiB per GPU, held until
 * the process ends), so it serves n <= 24 and returns EMU_SV_EINVAL beyond.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The first call on a device looks up the array's
 * address with that device current and clears its reduction ticket with a
 * fill kernel on `stream`.  Every call on a GPU shares that GPU's array and
 * it keeps the last call's contents: calls that overlap on one GPU (two
 * streams, or two threads) race.  Same arguments and results as
 * emu_sv_hpsi otherwise (the same kernels). */
int emu_sv_hpsi_static(int n, int dtype, const double *ham, double sigma_re,
                       double sigma_im, const void *v, double v_scale,
                       const void *prev, double prev_scale, void *w,
                       double *dots_out, int tile_log, int min_contig, int flags,
                       hipStream_t stream);

/* One fused Lanczos iteration step (orthogonalise, then the next matvec):
 *   z = w_scale*w - a*(pa_scale*pa) - b*(pb_scale*pb)   (stored back into w)
 *   r = sigma*H*z
 * norm2_out[0..4] = ||z||^2, <z|pb_s>, <z|w_s> (complex as re, im) with
 *   pb_s = pb_scale*pb and w_s = w_scale*w (the input before the update);
 * dots_out[0..2] = ||r||^2, <z|r>.
 * With H Hermitian, <v_j|sigma H v_{j+1}> = sigma conj(<z|w_s>)/(|z| conj(sigma))
 * for v_{j+1} = z/|z|, w_s = sigma H v_j, v_j = pb_s: so every Lanczos
 * coefficient of the next iteration (krylov_exp.py:54-61, two-vector MGS)
 * comes from this call, whose first H.psi pass performs the
 * orthogonalisation of krylov_exp.py:56-61 and whose passes are the matvec of
 * krylov_exp.py:50.  One call and one host read-back per Lanczos iteration.
 * w, r, pa, pb must be distinct; pa may be NULL. */
int emu_sv_lanczos_matvec(int n, int dtype, const double *ham, double sigma_re,
                          double sigma_im, void *w, double w_scale,
                          const void *pa, double pa_scale, double a_re,
                          double a_im, const void *pb, double pb_scale,
                          double b_re, double b_im, void *r, double *norm2_out,
                          double *dots_out, void *workspace, int tile_log,
                          int min_contig, int flags, hipStream_t stream);

/* emu_sv_lanczos_matvec with its coefficients in device memory, so that a
 * caller can enqueue iteration j+1 before reading iteration j's results:
 * coef = {Re, Im of a*pa_scale, Re, Im of b*pb_scale, pb_scale, w_scale}
 * (6 doubles, device), as written by emu_sv_lanczos_next_coef.  Tile or LDS
 * kernel kind (flags; the wave kind is rejected), n >= the tile sizes.  Same
 * outputs as emu_sv_lanczos_matvec (krylov_exp.py:50-61). */
int emu_sv_lanczos_matvec_dev(int n, int dtype, const double *ham, double sigma_re,
                              double sigma_im, void *w, const void *pa, const void *pb,
                              const double *coef, void *r, double *norm2_out,
                              double *dots_out, void *workspace, int tile_log,
                              int min_contig, int flags, hipStream_t stream);

/* emu_sv_lanczos_matvec_dev and emu_sv_lanczos_next_coef in one launch
 * sequence: the reduction that ends the iteration also writes next_coef (6
 * doubles, device; must not alias coef) and t_out (6 doubles) exactly as
 * emu_sv_lanczos_next_coef(norm2_out, dots_out, sigma, coef, next_coef,
 * t_out) would.  norm2_out, dots_out and t_out may be device memory or
 * pinned host memory, which that kernel then writes directly (so a caller
 * one iteration behind needs no copy command between iterations).
 * Replaces the same reference code as the two calls. */
int emu_sv_lanczos_iterate_dev(int n, int dtype, const double *ham, double sigma_re,
                               double sigma_im, void *w, const void *pa, const void *pb,
                               const double *coef, void *r, double *norm2_out,
                               double *dots_out, double *next_coef, double *t_out,
                               void *workspace, int tile_log, int min_contig, int flags,
                               hipStream_t stream);

/* On the device (one thread): the coefficients of the next fused Lanczos
 * iteration from the reductions of the current one (norm2 = the 5
 * orthogonalisation sums, dots = ||r||^2, <z|r>) and prev (the current
 * coefficient record, whose pb_scale becomes the next pa scale):
 *   n2 = ||z||, a = sigma conj(<z|w_s>)/(n2 conj(sigma)),
 *   b = <z|r>/n2^2 - a <z|pb_s>/n2        (krylov_exp.py:54-61 for j+1)
 * next = {a*s_j, b/n2, 1/n2, 1/n2}; t_out = {n2, b, a, ||r||^2/n2^2}. */
int emu_sv_lanczos_next_coef(const double *norm2, const double *dots, double sigma_re,
                             double sigma_im, const double *prev, double *next,
                             double *t_out, hipStream_t stream);

/* w <- w - a*(pa_scale*pa) - b*(pb_scale*pb); norm2_out[0] = ||w||^2.
 * The orthogonalisation of krylov_exp.py:54-61 (two-vector MGS, with the
 * coefficients computed by emu_sv_hpsi's fused dots). pa/pb may be NULL. */
int emu_sv_lanczos_update(int n, int dtype, void *w, const void *pa,
                          double pa_scale, double a_re, double a_im,
                          const void *pb, double pb_scale, double b_re,
                          double b_im, double *norm2_out, void *workspace,
                          hipStream_t stream);

/* out = sum_k coef_k * V_k for k < count (count <= 1024).  coef: host array of
 * 2*count doubles (re, im), already including vector scales.  V: host array
 * of device pointers.  out may alias V[0] (and no other V_k).  The final combination of
 * krylov_exp.py:66,88-90. */
int emu_sv_lincomb(int n, int dtype, int count, const void *const *vecs,
                   const double *coef, void *out, hipStream_t stream);

/* dots_out[0..2] = ||a||^2, Re<a|b>, Im<a|b> (b may be NULL).  Replaces the
 * vector_norm / vdot / tensordot calls of the reference on the path. */
int emu_sv_dot(int n, int dtype, const void *a, const void *b, double *dots_out,
               void *workspace, hipStream_t stream);

/* v <- s * v (real s). */
int emu_sv_scale(int n, int dtype, void *v, double s, hipStream_t stream);

/* Full modified Gram-Schmidt of w against count basis vectors V_k (stored
 * unnormalised; the orthonormal vector is scales[k] * V_k), in the order of
 * the Arnoldi branch of krylov_exp.py:52-58: for k = 0..count-1,
 * h_k = <v_k|w>, w <- w - h_k v_k.  The coefficients never leave the device:
 * launch k subtracts the previous projection (read from records_out) and
 * reduces the next overlap in the same pass over w (for 10 <= n <= 13 and
 * count <= 128, one workgroup holds w and runs the whole sequence in one
 * launch).  records_out (device,
 * 4*(count+1) doubles): record k < count = {||w||^2 before the k-th
 * subtraction, Re, Im <V_k|w>} (unscaled: h_k = scales[k] * <V_k|w>), record
 * count = {||w||^2 after the last subtraction}.  vecs: host array of device
 * pointers; scales: host doubles; count in [1, 1024]. */
int emu_sv_arnoldi_orthogonalize(int n, int dtype, int count, const void *const *vecs,
                                 const double *scales, void *w, double *records_out,
                                 void *workspace, hipStream_t stream);

/* emu_sv_arnoldi_orthogonalize, then w <- w / ||w|| with the norm after the
 * last subtraction (records_out[4*count], still the unnormalised value), on
 * the device: w becomes the next orthonormal Arnoldi vector v_{j+1} =
 * w / T[j+1, j] (krylov_exp.py:60-61,74) without a host round trip, so the
 * next iteration can be enqueued before the host has read this one.  A zero
 * norm (exact breakdown) leaves w as it is. */
int emu_sv_arnoldi_orthonormalize(int n, int dtype, int count, const void *const *vecs,
                                  const double *scales, void *w, double *records_out,
                                  void *workspace, hipStream_t stream);

/* out = exp(sigma H) v: one Krylov time step of the Hermitian branch of
 * krylov_exp (emu_base/math/krylov_exp.py:23-131 with is_hermitian=True, as
 * called by emu_sv/time_evolution/torch.py:161-169) with its whole loop in
 * the library: the opening H.psi with its dots, then per iteration
 * emu_sv_lanczos_iterate_dev (the next iteration's coefficients formed on
 * the device) and an asynchronous copy of its records; the host runs the
 * reference's Expokit test one iteration behind (emu_sv_lanczos_exp on the
 * tridiagonal T) and enqueues iteration j+1 before waiting for iteration j
 * unless half the last error estimate is below the tolerance; the happy
 * breakdown and the final combination of m+1 basis vectors (emu_sv_lincomb)
 * follow the reference.  Waits for each iteration's records on the host;
 * the final combination is left enqueued on stream.
 *   v          device input state (2^n elements, not modified; out may alias it)
 *   basis      device buffer of max_dim + 1 vectors of 2^n elements
 *   records    device, emu_sv_krylov_step_records_doubles(max_dim) doubles
 *   host_records  pinned host buffer of the same size; the reduction that
 *                 ends each iteration writes the iteration's record into it
 *                 directly (pageable memory also works, through a copy per
 *                 iteration)
 *   flags      kernel kind (tile, btile or LDS; n must cover its tiles),
 *              EMU_SV_FLAG_COMPLEX
 *   result_out host int[4]: iterations, converged, happy breakdown, matvecs
 *   decomp_out host, emu_sv_krylov_step_decomp_doubles(max_dim) doubles, or
 *              NULL: the Lanczos decomposition for callers that reuse the
 *              Krylov space (the adjoint).  With M = max_dim + 2 and m the
 *              iterations: [0] m, [1] T[m, m-1] (0 after a happy breakdown),
 *              [2] happy, [3] converged; T[k, k] (complex) at 4 + 2k;
 *              T[k+1, k] (real) at 4 + 2M + k; T[k, k+1] (complex) at
 *              4 + 3M + 2k; the basis scales (basis vector k is scale_k times
 *              v for k = 0, basis row k for k >= 1) at 4 + 5M + k; the
 *              combination coefficients exp(T_ext) e_0 (complex, m + 1 of
 *              them, m after a breakdown) at 4 + 6M + 2k.
 * max_dim in [1, 256].  Returns EMU_SV_OK also when the iteration did not
 * converge (result_out[1] == 0: the reference's RecursionError). */
size_t emu_sv_krylov_step_records_doubles(int max_dim);
size_t emu_sv_krylov_step_decomp_doubles(int max_dim);
int emu_sv_krylov_step(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                       const void *v, void *out, void *basis, int max_dim, double exp_tolerance,
                       double norm_tolerance, double *records, double *host_records,
                       void *workspace, int tile_log, int min_contig, int flags,
                       hipStream_t stream, int *result_out, double *decomp_out);

/* The Arnoldi Gram-Schmidt step in its Gram-matrix form, for the pipelined
 * engines on vectors outside the one-workgroup range.  Replaces the same
 * loop as emu_sv_arnoldi_orthogonalize (emu_base/math/krylov_exp.py:54-61
 * with is_hermitian=False): the reference's modified Gram-Schmidt
 * coefficients h_k = <a_k|w_k> equal c_k - sum_{i<k} h_i G(k, i) with
 * c_k = <a_k|w> and G(k, i) = <a_k|a_i>, so one sweep of overlaps (which
 * also reduces the new Gram row of a_j) and one subtraction sweep replace
 * the 2 (count + 1) launches of the chain: 2 count + 3 vector transfers
 * instead of 4 count + 6.  The basis is kept unnormalised, a_k = s_k u_k,
 * with the scales on the device, so the next matvec can be enqueued on
 * u_{j+1} with scale 1 before this call's records are read.
 *   vecs     u_0 .. u_j (count = j + 1 <= 256 device vectors of 2^n elements)
 *   w        in: op(u_j) (unscaled); out: u_{j+1} = s_j w - sum_k h_k a_k
 *   v0_scale s_0 (1 / ||u_0|| for the Krylov start vector)
 *   state    device doubles, emu_sv_arnoldi_gram_state_doubles(ld): the
 *            scales, the Gram matrix of the basis and scratch; ld > count.
 *            Needs no clearing: every entry a call reads was written by an
 *            earlier call of the same run (or by this one)
 *   records_out  device (count + 1) x 4 doubles: row k < count =
 *            [0, Re h_k, Im h_k, 0] (the Hessenberg column T[k, j]), with
 *            row 0 [0] = ||s_j w||^2; row count = [||u_{j+1}||^2, 0, 0, 0]
 *            (T[j+1, j] is its root; s_{j+1} = 1 / root is kept in state,
 *            0 when it is zero).
 * Calls of one run go in order count = 1, 2, ... on one stream. */
size_t emu_sv_arnoldi_gram_state_doubles(int ld);
int emu_sv_arnoldi_gram(int n, int dtype, int count, const void *const *vecs, void *w,
                        double v0_scale, double *state, int ld, double *records_out,
                        hipStream_t stream);

/* out = exp(sigma H) v by the Arnoldi branch of krylov_exp
 * (emu_base/math/krylov_exp.py:23-107 with is_hermitian=False) with its loop
 * inside the library: per iteration emu_sv_hpsi (flags: kernel kind,
 * EMU_SV_FLAG_COMPLEX, EMU_SV_FLAG_NOISE; KEEP_DIAG is set by the driver
 * after the first matvec), emu_sv_arnoldi_orthonormalize and an
 * asynchronous copy of the records; the host's column exponential
 * (emu_sv_expm_column), the Expokit estimate, the happy-breakdown test and
 * the final combination (emu_sv_lincomb) follow the reference.  Iteration
 * j+1 is enqueued before the host waits for iteration j, unless half the
 * last error estimate is below the tolerance; a discarded speculative
 * iteration costs one matvec.  Synchronises the stream.
 *   v          device input state (2^n elements, not modified; out may alias it)
 *   basis      device buffer of max_dim vectors of 2^n elements (v_1 .. v_max_dim)
 *   records    device, emu_sv_arnoldi_exp_records_doubles(max_dim) doubles
 *              (the per-iteration records, then the Gram-Schmidt state)
 *   host_records  pinned host buffer of the same size
 * The Gram-Schmidt step is the one-workgroup emu_sv_arnoldi_orthonormalize
 * for 10 <= n <= 13 (max_dim <= 128), else emu_sv_arnoldi_gram.
 *   result_out host int[4]: iterations, converged, happy breakdown, matvecs
 * max_dim in [1, 256].  Returns EMU_SV_OK also when the iteration did not
 * converge (result_out[1] == 0: the reference's RecursionError). */
size_t emu_sv_arnoldi_exp_records_doubles(int max_dim);
int emu_sv_arnoldi_exp(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                       void *v, void *out, void *basis, int max_dim, double exp_tolerance,
                       double norm_tolerance, double *records, double *host_records,
                       void *workspace, int tile_log, int min_contig, int flags,
                       hipStream_t stream, int *result_out);

/* psi <- (A acting on qubit bit b) psi, A a 2x2 complex matrix (8 doubles
 * row-major (re, im), host) in the (bit 0, bit 1) = (g, r) basis: for each
 * pair (x, x | 2^b) with bit b clear, (a0, a1) <- (A00 a0 + A01 a1,
 * A10 a0 + A11 a1).  Replaces emu_sv/algebra.py:28-37 (the quantum-jump
 * operator, qubit q = bit n-1-q). */
int emu_sv_apply_1q(int n, int dtype, void *psi, int bit, const double *a,
                    hipStream_t stream);

/* out (+)= S acting on the bit pair (bit_a, bit_b): S a 4x4 complex matrix
 * (32 doubles row-major (re, im), host) over the local index
 * 2*x_{bit_a} + x_{bit_b}; accumulate != 0 adds to out.  in may alias out.
 * On a density matrix stored as a 2n-bit vector (row bits high), the pair
 * (row bit, column bit) of one qubit carries K (x) I and L (x) conj(L), the
 * jump terms L rho L^+ of the Lindbladian (emu_sv/lindblad_operator.py:94-135,
 * 203-213). */
int emu_sv_apply_pair(int n, int dtype, const void *in, void *out, int bit_a,
                      int bit_b, const double *op, int accumulate, hipStream_t stream);

/* out += sum over the qubits j < n_side of S on the bit pair (j + n_side, j)
 * of in: the jump and noise terms of every qubit of an n_side-qubit density
 * matrix (2*n_side-bit vector, row bits high) in one launch, in the gathered
 * form (each output element reads its 4 n_side partners).  For small
 * matrices, n_side <= 8; in and out must not alias.  Same terms as n_side
 * calls of emu_sv_apply_pair (emu_sv/lindblad_operator.py:94-135, 203-213). */
int emu_sv_apply_pairs_all(int n_side, int dtype, const void *in, void *out,
                           const double *op, hipStream_t stream);

/* out = s (M - M^H) for a 2^side_log x 2^side_log row-major complex matrix M
 * (s complex); out may alias M.  The anti-Hermitian part taken twice in
 * RydbergLindbladian.__matmul__ (emu_sv/lindblad_operator.py:196-216). */
int emu_sv_antihermitian(int side_log, int dtype, const void *m, void *out,
                         double s_re, double s_im, hipStream_t stream);

/* Bytes of device workspace emu_sv_sample needs for 2^n categories. */
size_t emu_sv_sample_workspace_bytes(int n);

/* Inverse-CDF sampling of basis indices: for each of count draws u_k in
 * [0, 1) (uniforms, device doubles), indices_out[k] (device int64) = the first
 * x whose cumulative probability exceeds u_k * sum(p), clamped to 2^n - 1
 * (torch.searchsorted(cdf, u, right=True) of emu_sv/utils.py:17-49).
 * p(x) = |src[x]|^2 (mode 0, a state vector) or |src[x * stride]| (mode 1,
 * the diagonal of a density matrix with stride 2^n + 1).  FP64 sums; the
 * draws themselves come from the caller.  Replaces the multinomial /
 * inverse-CDF sampling of DenseState.sample (emu_sv/dense_state.py:40-95). */
int emu_sv_sample(int n, int dtype, const void *src, uint64_t stride, int mode,
                  const double *uniforms, int count, int64_t *indices_out,
                  void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Copy with element-type conversion (c128 <-> c64 or same type). */
int emu_sv_convert(int n, int src_dtype, const void *src, int dst_dtype,
                   void *dst, hipStream_t stream);

/* Density-density correlation <n_i n_j>, n x n float64 into corr_out
 * (device), qubit-indexed; its diagonal is the occupation <n_i>.
 * Replaces qubit_occupation_sv_impl and correlation_matrix_sv_impl
 * (emu_sv/custom_callback_implementations.py:30-47, 77-102). */
int emu_sv_correlation(int n, int dtype, const void *psi, double *corr_out,
                       void *workspace, hipStream_t stream);

/* Parameter traces of a sum of bilinear forms, the per-qubit sums behind
 * every Hamiltonian-parameter gradient (the DHDOmegaSparse / DHDDeltaSparse /
 * DHDPhiSparse / DHDUSparse traces of emu_sv/time_evolution/torch.py:8-119,
 * used by its backward at :301-366).  For count pairs of device vectors
 * W_l, Z_l (row l at element offset l*stride) and every BIT position b:
 *   up[b]  = sum_l sum_{x: x_b=1} conj(W_l(x)) Z_l(x ^ 2^b)   (<W|s+_b|Z>)
 *   dn[b]  = sum_l sum_{x: x_b=0} conj(W_l(x)) Z_l(x ^ 2^b)   (<W|s-_b|Z>)
 *   occ[b] = sum_l sum_{x: x_b=1} conj(W_l(x)) Z_l(x)         (<W|n_b|Z>)
 *   pair[b][b'] (b < b', only if want_pairs)
 *          = sum_l sum_{x: x_b=x_b'=1} conj(W_l(x)) Z_l(x)    (<W|n_b n_b'|Z>)
 * out (device, complex interleaved): up[n], dn[n], occ[n], pair[n*n]
 * (6n + 2n^2 doubles; the pair block is written only if want_pairs).
 * Deterministic (per-workgroup partials, fixed-order sums). */
int emu_sv_parameter_traces(int n, int dtype, int count, const void *w,
                            const void *z, uint64_t stride, int want_pairs,
                            double *out, void *workspace, hipStream_t stream);

/* ---- many registers in one device-driven engine ------------------------ */
/* One register of a batch.  All pointers are device memory owned by the
 * caller and 16-byte aligned.  psi holds the initial state and receives the
 * final state. */
#define EMU_SV_BATCH_TILE_LOG(t) (((t) & 0xF) << 16)
#define EMU_SV_BATCH_MIN_CONTIG(c) (((c) & 0xF) << 20)
typedef struct emu_sv_batch_register {
  int n;              /* qubits (1..30) */
  int steps;          /* time steps */
  int flags;          /* EMU_SV_FLAG_COMPLEX if any step has phi != 0; optionally */
                      /* this register's own H.psi schedule:                   */
                      /* EMU_SV_BATCH_TILE_LOG(t) | EMU_SV_BATCH_MIN_CONTIG(c)  */
                      /* (0: the batch's tile_log / min_contig)                */
  int pause_each_step; /* 1: stop after every step until emu_sv_batch_resume */
  const double *ham;  /* steps * EMU_SV_HAM_DOUBLES packed blocks, one per step */
  const double *dt;   /* steps time steps in us (op = -i dt H) */
  void *psi;          /* 2^n elements: initial state in, final state out */
  void *psi_alt;      /* 2^n elements of scratch */
  void *rows;         /* (max_dim + 1) * 2^n elements of scratch (Krylov basis) */
} emu_sv_batch_register;

/* Bytes of device workspace emu_sv_batch_create needs for these registers. */
size_t emu_sv_batch_workspace_bytes(int count, int dtype, const emu_sv_batch_register *regs,
                                    int tile_log, int min_contig);

/* Set up a batch: every register runs the Krylov time stepping of
 * emu_base/math/krylov_exp.py:23-131 (the arithmetic of emu_sv_lanczos_matvec,
 * the Expokit test of emu_sv_lanczos_exp) step after step
 * (emu_sv/sv_backend_impl.py:186-209), with its whole state machine in device
 * memory, so one batched iteration is a fixed kernel sequence (captured once
 * in a hipGraph).  max_dim <= 62; tile_log 8..13 (complex64: 8..14) /
 * min_contig: the H.psi pass
 * split (0 = default).  The library allocates no device memory: the
 * workspace (256-byte aligned, zero-filled not required) is the caller's.  *handle_out is a
 * host object released by emu_sv_batch_destroy.  No counterpart in the
 * reference (it evolves one register per call). */
int emu_sv_batch_create(int count, int dtype, const emu_sv_batch_register *regs, int max_dim,
                        double tolerance, int tile_log, int min_contig, void *workspace,
                        size_t workspace_bytes, hipStream_t stream, void **handle_out);

/* Replay the batched iteration up to max_iterations times on stream, checking
 * every check_every iterations how many registers are still running
 * (*remaining_out; paused, done and failed registers are not running).
 * Registers whose Krylov iteration did not converge within max_dim stop
 * (phase 4 in emu_sv_batch_status). */
int emu_sv_batch_run(void *handle, int max_iterations, int check_every, hipStream_t stream,
                     int *remaining_out);

/* Copy the final state of every done (or failed) register into its psi
 * (the state may sit in psi_alt after an odd number of steps). */
int emu_sv_batch_publish(void *handle, hipStream_t stream);

/* Per register (host arrays of count ints, any may be NULL): phase (0 start,
 * 1 iterating, 2 combining, 3 done, 4 failed), completed steps, matvecs. */
int emu_sv_batch_status(void *handle, int *phase_out, int *steps_out, int *matvecs_out,
                        hipStream_t stream);

/* Restart every register paused after a step (pause_each_step). */
int emu_sv_batch_resume(void *handle, hipStream_t stream);

/* The Lanczos decomposition of register reg's last finished step (valid
 * while it is paused, or done): info_out[0..4] = phase, completed steps, m
 * (combination coefficients / basis vectors used), 1 if the step ended in a
 * happy breakdown, matvecs; scalars_out[0] = the norm of the step's initial
 * state; arrays_out (host, emu_sv_batch_array_layout()[5] doubles) = T[k,k],
 * T[k+1,k], the basis scales, the coefficients and T[k,k+1]; state_out = the
 * current state, prev_state_out = the step's initial state (basis vector 0),
 * rows_out = basis vectors 1.. (device pointers). */
int emu_sv_batch_step_info(void *handle, int reg, int *info_out, double *scalars_out,
                           double *arrays_out, void **state_out, void **prev_state_out,
                           void **rows_out, hipStream_t stream);

/* Offsets (doubles) of the arrays in emu_sv_batch_step_info's arrays_out:
 * [T diag, T sub, scales, coefficients, T super, total]. */
int emu_sv_batch_array_layout(int *offsets_out);

int emu_sv_batch_destroy(void *handle);

/* Pack steps parameter blocks (omega/delta/phi: steps x n, row-major;
 * interaction n x n) into blocks_out (steps * EMU_SV_HAM_DOUBLES, host). */
int emu_sv_pack_hamiltonian_steps(int n, int steps, const double *omega, const double *delta,
                                  const double *phi, const double *interaction, double *blocks_out);

/* Host: E = expm(A) for an m x m complex128 matrix (row-major, interleaved),
 * scaling and squaring with a [13/13] Pade approximant (Higham 2005).
 * Replaces torch.linalg.matrix_exp on the (j+3)x(j+3) extended Krylov
 * matrix (emu_base/math/krylov_exp.py:65,78). */
int emu_sv_expm(int m, const double *a, double *e);

/* Host: col = expm(A) e_0, the first column only (m complex, interleaved),
 * by diagonal shift and Taylor substeps of norm <= 4 (the Pade form of
 * emu_sv_expm beyond ||A - cI||_inf = 256).  Every use of the extended
 * Krylov matrix's exponential reads column 0 (krylov_exp.py:65-90). */
int emu_sv_expm_column(int m, const double *a, double *col_out);

/* Host: the exponential of a Hermitian Lanczos matrix from its symmetric
 * tridiagonal form.  tdiag: m complex T[k,k] (interleaved), tsub: m-1 real
 * T[k+1,k]; sigma: the operator scale (op = sigma*H); n2: the newest
 * T[m,m-1] (the norm of the next Lanczos vector).  Writes
 *   out[0] = |expd[m,0]|, out[1] = |expd[m+1,0]|   (expd = exp of the
 *            Expokit extended matrix, the err1 / err2/n terms of
 *            emu_base/math/krylov_exp.py:81-84), and if want_coeffs,
 *   out[2 .. 2+2m) = exp(T) e_0 and out[2+2m .. 4+2m) = expd[m,0] (complex,
 *            interleaved): the combination coefficients of krylov_exp.py:88.
 * Replaces torch.linalg.matrix_exp of the (j+3)x(j+3) extended matrix
 * (krylov_exp.py:77-78) for the Hermitian Lanczos path: O(m^2) per call
 * (O(m^3) with coefficients) instead of O(m^3). */
int emu_sv_lanczos_exp(int m, const double *tdiag, const double *tsub,
                       double sigma_re, double sigma_im, double n2,
                       int want_coeffs, double *out);

/* Host: eigen-decomposition J = Z diag(lambda) Z^T of the real symmetric
 * tridiagonal J (diagonal d, m values; off-diagonal e, m-1 values) by the
 * implicit QR iteration of emu_sv_lanczos_exp; z_out is m x m row-major
 * (row k = component k of every eigenvector).  Used for the Frechet weights
 * of the adjoint (the integrals of products of two Krylov trajectories)
 * in place of a block matrix exponential (double_krylov.py:18-58). */
int emu_sv_tridiag_eig(int m, const double *d, const double *e, double *lambda_out,
                       double *z_out);

#ifdef __cplusplus
}
#endif
#endif /* EMU_SV_HIP_H */
