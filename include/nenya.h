/*
 * nenya.h -- C ABI of the MI355X (gfx950) Nenya prover.
 *
 * Drop-in replacement for the reference prover API (tnyuzg/nenya,
 * crates/nenya/snark-ring/src/lib.rs:23-36): R1CS instance + witness in,
 * serialized proof bytes out, same BLAKE3 Fiat-Shamir transcript, so the
 * reference verifier accepts the bytes.
 *
 * Conventions
 *   - k in {8, 16, 32, 64}: the base ring Z/2^k.  Every limb buffer
 *     (`vals`, witnesses) holds one little-endian k-bit integer per element,
 *     i.e. arrays of uint8_t / uint16_t / uint32_t / uint64_t.
 *   - Sparse matrices are COO triples (rows[i], cols[i], vals[i]), i < nnz;
 *     duplicates sum (reduction/src/matrix.rs:11-18).
 *   - Return 0 on success or a NENYA_E_* code; nenya_last_error() describes
 *     the failure (thread-local).  Nothing panics or unwinds across the ABI
 *     (the reference asserts on shape misuse: simd.rs:224-231,
 *     reduce.rs:463-467).
 *   - Proof buffers are library-allocated; release them with nenya_free().
 *     An owned proof is the pinned host buffer it was assembled in, handed
 *     over without a copy; nenya_free keeps up to four released buffers for
 *     later proofs, so a caller that frees each proof before the next one
 *     allocates nothing in steady state.
 *   - A params object is immutable after creation and owns its device memory
 *     and HIP stream; calls on one params object are serialized internally.
 *   - Threading: a single-proof call (nenya_prove_*) that finds no other
 *     library call running and no server alive pipelines its sumcheck rounds:
 *     its kernels wait for challenges its thread posts.  Every other library
 *     call waits until it is done, and it waits for calls already running
 *     (then it runs without pipelining), so library calls from several
 *     threads are safe.  Other HIP runtime calls the caller makes from other
 *     threads meanwhile (allocation, frees, synchronization) can stall such a
 *     proof until its device waits time out (10 s, NENYA_E_DEVICE; never a
 *     wrong proof): drive the device from one thread, or use a server.
 */
#ifndef NENYA_H
#define NENYA_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NENYA_OK 0
#define NENYA_E_SHAPE 1  /* a shape/size mismatch (the reference's assert!) */
#define NENYA_E_OOM 2    /* device memory exhausted */
#define NENYA_E_DEVICE 3 /* HIP runtime error or internal self-check */
#define NENYA_E_ARG 4    /* invalid argument (width, null pointer, ...) */
#define NENYA_E_NOGPU 5  /* no HIP device visible */
#define NENYA_E_CANCELED 6 /* a served job canceled with nenya_cancel */

typedef struct nenya_params nenya_params;

/* SimdSnarkParams::secure(log_n, log_m, seed_w, seed_q)
 * (snark-ring/src/simd.rs:112-119): Pcs::setup for the packed witness and
 * R2Pcs::setup for the kernel error at delta=0.19, lambda=100, rate 1/4. */
int nenya_params_simd_secure(int k, int log_n, int log_m, uint64_t seed_w, uint64_t seed_q, nenya_params** out);

/* prove_simd(&SimdR1csInstance, &params) (snark-ring/src/simd.rs:212-217):
 * A, B, C are n x m (n = 2^log_n, m = 2^log_m); `witnesses` holds the 32
 * witnesses back to back (32 * m limbs).  Returns SimdSnarkProof::serialize. */
int nenya_prove_simd(const nenya_params* p,
                     const uint32_t* a_rows, const uint32_t* a_cols, const void* a_vals, size_t a_nnz,
                     const uint32_t* b_rows, const uint32_t* b_cols, const void* b_vals, size_t b_nnz,
                     const uint32_t* c_rows, const uint32_t* c_cols, const void* c_vals, size_t c_nnz,
                     const void* witnesses, uint8_t** proof, size_t* proof_len);

/* Same as nenya_prove_simd, but *proof points into a pinned host buffer owned
 * by `p` (no copy, nothing allocated per call); the bytes stay valid until
 * the next proof made with `p` or nenya_params_free(p).  No reference
 * counterpart (the reference returns an owned proof): for callers that read
 * the proof once, e.g. to send it, without keeping a copy. */
int nenya_prove_simd_view(const nenya_params* p,
                          const uint32_t* a_rows, const uint32_t* a_cols, const void* a_vals, size_t a_nnz,
                          const uint32_t* b_rows, const uint32_t* b_cols, const void* b_vals, size_t b_nnz,
                          const uint32_t* c_rows, const uint32_t* c_cols, const void* c_vals, size_t c_nnz,
                          const void* witnesses, const uint8_t** proof, size_t* proof_len);

/* MimdSnarkParams::secure(log_n, log_m, pack_mimd_matrices(a, b, c), seed_w,
 * seed_e) (snark-ring/src/mimd.rs:84-107, reduction/src/reduce.rs:249-273):
 * 32 independent n x m instances (n = 2^log_n, m = 2^log_m).  Each matrix
 * set is the 32 instance matrices in COO concatenated, instance s holding
 * entries off[s] .. off[s+1]-1 (33 offsets); the union-support packing is
 * done here, once, on setup. */
int nenya_params_mimd_secure(int k, int log_n, int log_m,
                             const size_t* a_off, const uint32_t* a_rows, const uint32_t* a_cols, const void* a_vals,
                             const size_t* b_off, const uint32_t* b_rows, const uint32_t* b_cols, const void* b_vals,
                             const size_t* c_off, const uint32_t* c_rows, const uint32_t* c_cols, const void* c_vals,
                             uint64_t seed_w, uint64_t seed_e, nenya_params** out);

/* prove_mimd(&witnesses, &params) (snark-ring/src/mimd.rs:235-240):
 * `witnesses` holds the 32 witnesses back to back (32 * m limbs).  Returns
 * MimdSnarkProof::serialize. */
int nenya_prove_mimd(const nenya_params* p, const void* witnesses, uint8_t** proof, size_t* proof_len);

/* View variant of nenya_prove_mimd (as nenya_prove_simd_view; no reference
 * counterpart). */
int nenya_prove_mimd_view(const nenya_params* p, const void* witnesses, const uint8_t** proof, size_t* proof_len);

/* GeneralSnarkParams::secure(log_blocks, pack_general_matrices(a, b, c),
 * seed_w, seed_e) (snark-ring/src/general.rs:69-95,
 * reduction/src/reduce.rs:618-651): the cyclic-diagonal packing of the
 * square n x n matrices (n = 32 * 2^log_blocks) is done here, once, on setup. */
int nenya_params_general_secure(int k, int log_blocks,
                                const uint32_t* a_rows, const uint32_t* a_cols, const void* a_vals, size_t a_nnz,
                                const uint32_t* b_rows, const uint32_t* b_cols, const void* b_vals, size_t b_nnz,
                                const uint32_t* c_rows, const uint32_t* c_cols, const void* c_vals, size_t c_nnz,
                                uint64_t seed_w, uint64_t seed_e, nenya_params** out);

/* prove_general(w, &params) (snark-ring/src/general.rs:367-372): w has
 * n = 32 * 2^log_blocks limbs in host memory.  Returns
 * GeneralSnarkProof::serialize. */
int nenya_prove_general(const nenya_params* p, const void* w, size_t n, uint8_t** proof, size_t* proof_len);

/* View variant of nenya_prove_general (as nenya_prove_simd_view; no
 * reference counterpart). */
int nenya_prove_general_view(const nenya_params* p, const void* w, size_t n, const uint8_t** proof,
                             size_t* proof_len);

/* Same as nenya_prove_general with `w` already resident in device memory
 * (a hipMalloc'd pointer on the current device).  No reference counterpart:
 * lets a caller that produced the witness on the GPU skip the host copy. */
int nenya_prove_general_device(const nenya_params* p, const void* d_w, size_t n, uint8_t** proof, size_t* proof_len);

/* Same as nenya_prove_general_device, but *proof points into a pinned host
 * buffer owned by `p` (no copy, nothing allocated per call); the bytes stay
 * valid until the next proof made with `p` or nenya_params_free(p). */
int nenya_prove_general_device_view(const nenya_params* p, const void* d_w, size_t n, const uint8_t** proof,
                                    size_t* proof_len);

/* Same proof as nenya_prove_general_device, for callers that cannot allocate
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The call allocates no
 * device or pinned memory and creates no stream or event: every device
 * buffer of the proof is carved out of a static 256 MiB device array (one
 * per GPU, held until the process ends), the pinned staging areas and proof
 * buffer were sized when `p` was created, and the proof is copied into the
 * caller's `proof` buffer of `proof_cap` bytes.  Limit: log_blocks <= 10
 * (n <= 2^15 constraints); larger params get NENYA_E_SHAPE, a buffer smaller
 * than the proof gets NENYA_E_ARG with the needed size in *proof_len.  Calls
 * on one GPU share its array (whose contents the last call leaves behind),
 * so the library runs them one at a time. */
int nenya_prove_general_device_noalloc(const nenya_params* p, const void* d_w, size_t n, uint8_t* proof,
                                       size_t proof_cap, size_t* proof_len);

/* Per-stage timing (ms) of the last proof made with `p` when enabled
 * (GeneralProveTiming / SimdProveTiming, general.rs:153-173,
 * simd.rs:177-192): HIP events on the prover stream at stage boundaries, no
 * added synchronization, so a stage is the time the prover stream took to
 * get from one boundary to the next in the overlapped schedule.
 * nenya_last_timing writes up to `cap` entries and returns the count; names
 * are separated by '\n' in `names` (up to names_cap bytes). */
int nenya_set_timing(nenya_params* p, int enabled);
int nenya_last_timing(const nenya_params* p, double* ms, int cap, char* names, size_t names_cap);
/* Wall-clock timing (ms) of the stages of the setup that made `p`
 * (context: the params' stream, pool and device constants; pcs_params: the
 * PCS parameters with the RAA permutations and diagonals,
 * raa.rs:43-63; pcs_upload; matrices: the matrix packing of
 * pack_general_matrices / pack_mimd_matrices, reduce.rs:249-273, 372-438,
 * 618-651; staging), in the same format as nenya_last_timing (no reference
 * counterpart: the reference does not time its setup). */
int nenya_setup_timing(const nenya_params* p, double* ms, int cap, char* names, size_t names_cap);

/* Per-kernel device timing for roofline figures (no reference counterpart):
 * when enabled, every launch on the params' stream is bracketed by HIP
 * events.  nenya_kernel_stats returns, per kernel name, the summed duration
 * (ms), launch count, algorithmic bytes and limb multiply-accumulates over
 * the proofs made since profiling was (re)enabled. */
int nenya_set_profile(nenya_params* p, int enabled);
/* Restrict profiling to launches of one kernel name (NULL or "" = all), so a
 * timed run can time its dominant kernel with negligible event overhead. */
int nenya_set_profile_only(nenya_params* p, const char* kernel);
int nenya_kernel_stats(const nenya_params* p, int cap, char* names, size_t names_cap, double* ms, double* count,
                       double* bytes, double* macs);

/* ---- Verifiers: verify_simd / verify_mimd / verify_general
 * (snark-ring/src/simd.rs:380-522, mimd.rs:403-551, general.rs:545-698) on
 * serialized proofs.  Transcript, sumcheck and PCS checks run on the host;
 * the linear-size row recompute runs on the GPU with the params' matrices.
 * The call returns NENYA_OK or an error code; *verdict receives 0 (Ok) or the
 * VerifyError variant in error.rs:5-16 order: 1 MultSumcheck, 2 LinSumcheck,
 * 3 LinClaimMismatch, 4 LinARowMismatch, 5 PcsW, 6 PcsE, 7 KernelCheck; and 8
 * when the bytes do not parse as a proof of these params (the reference
 * verifies proof structs and has no such case). */
int nenya_verify_simd(const nenya_params* p,
                      const uint32_t* a_rows, const uint32_t* a_cols, const void* a_vals, size_t a_nnz,
                      const uint32_t* b_rows, const uint32_t* b_cols, const void* b_vals, size_t b_nnz,
                      const uint32_t* c_rows, const uint32_t* c_cols, const void* c_vals, size_t c_nnz,
                      const uint8_t* proof, size_t proof_len, int* verdict);
int nenya_verify_mimd(const nenya_params* p, const uint8_t* proof, size_t proof_len, int* verdict);
int nenya_verify_general(const nenya_params* p, const uint8_t* proof, size_t proof_len, int* verdict);

/* ---- Proof container (no reference counterpart: the reference only
 * serializes proof structs).  A 72-byte versioned header in front of the
 * serialized proof: magic "NNYA", version 1, model (0 SIMD, 1 General,
 * 2 MIMD), k, log_n / log_blocks, log_m, the two PCS seeds, the payload
 * length and a BLAKE3 digest of the payload.  nenya_proof_container wraps
 * the proof made with `p` (output freed with nenya_free);
 * nenya_proof_payload checks a container's header and digest and points at
 * its payload (NENYA_E_ARG when it is not an intact container).  The
 * nenya_verify_* calls accept a container in place of the raw bytes; one
 * made for params of another model, width, shape or seeds, or with a wrong
 * digest, gets verdict 8. */
int nenya_proof_container(const nenya_params* p, const uint8_t* proof, size_t len, uint8_t** out, size_t* out_len);
int nenya_proof_payload(const uint8_t* c, size_t len, const uint8_t** payload, size_t* payload_len, int* model,
                        int* k);

/* ---- Batched and served proofs (no reference counterpart: the reference
 * proves one instance per call, general.rs:367-372, simd.rs:212-217,
 * mimd.rs:235-240).
 * A server owns `workers` host threads, each with its own HIP stream, device
 * pool and pinned proof buffers, so independent proofs (any params, any
 * width) overlap on one GPU.  Jobs start in submission order, each once its
 * device footprint fits `mem_budget` bytes (0 = 80% of the free HBM).  Input
 * buffers are borrowed until the job is done; params must outlive their
 * jobs.  nenya_wait returns the job's status and a view of its proof bytes,
 * valid until nenya_job_release; queued/run times are in milliseconds. */
typedef struct nenya_server nenya_server;
int nenya_server_create(int workers, size_t mem_budget, nenya_server** out);
int nenya_submit_general(nenya_server* s, const nenya_params* p, const void* w, size_t n, uint64_t* job);
int nenya_submit_simd(nenya_server* s, const nenya_params* p, const uint32_t* ar, const uint32_t* ac, const void* av,
                      size_t an, const uint32_t* br, const uint32_t* bc, const void* bv, size_t bn, const uint32_t* cr,
                      const uint32_t* cc, const void* cv, size_t cn, const void* witnesses, uint64_t* job);
int nenya_submit_mimd(nenya_server* s, const nenya_params* p, const void* witnesses, uint64_t* job);
/* 1 = done, 0 = queued or running, < 0 = -(error code) of the call itself */
int nenya_poll(nenya_server* s, uint64_t job);
int nenya_wait(nenya_server* s, uint64_t job, const uint8_t** proof, size_t* proof_len, double* queued_ms,
               double* run_ms);
int nenya_job_release(nenya_server* s, uint64_t job);
/* Cancels a job: a queued job is dropped before it starts; a running one
 * stops at its next checkpoint (each sumcheck round, each commitment and
 * opening), its kernels ending normally.  nenya_wait on it then returns
 * NENYA_E_CANCELED; other jobs are unaffected.  NENYA_E_ARG when the job is
 * unknown or already done.  A job whose device footprint exceeds the
 * server's budget is refused at submit with NENYA_E_OOM.  Test switch:
 * NENYA_INJECT_OOM=<job>:<n>, read when a server is created, makes the n-th
 * device allocation of that job (job ids count from 1) fail with
 * NENYA_E_OOM.  NENYA_DEVICE_FS=1, read when a server is created, makes its
 * workers draw every sumcheck round challenge on the device (the
 * transcript's BLAKE3 state and the ChaCha12 sampler run in the kernel that
 * lifts the challenge), so rounds need no host round trip; same bytes. */
int nenya_cancel(nenya_server* s, uint64_t job);
void nenya_server_free(nenya_server* s);

/* Autotuner (no reference counterpart): for the widths in k_mask (bit 0 k=8,
 * bit 1 k=16, bit 2 k=32, bit 3 k=64) and every work size 2^log_min ..
 * 2^log_max elements (8 <= log_min <= log_max <= 22), times the candidate
 * launch shapes of the sumcheck folds (k_fold_mx split or bulk mode and its
 * workgroup count; k_fold's workgroup cap at k=64) and of the linear-check
 * Gram sums (k_gram_lin workgroups) on synthetic tables, and makes the
 * fastest the shape every later launch of the process uses.  The picks,
 * with their times and the built-in choices' times, go to `report` as JSON
 * (up to report_cap bytes). */
int nenya_autotune(int k_mask, int log_min, int log_max, char* report, size_t report_cap);

/* Device-to-device copy of `bytes` bytes by the library's copy kernel
 * (2048 workgroups at most), synchronized.  No reference counterpart: a
 * transfer of known size for calibrating hardware traffic counters. */
int nenya_device_copy(void* dst, const void* src, size_t bytes);

void nenya_params_free(nenya_params* p);
void nenya_free(void* ptr);
const char* nenya_last_error(void);
int nenya_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* NENYA_H */
