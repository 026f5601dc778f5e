/*
 * sdim_hip.h - C ABI of the MI355X (gfx950) prime-dimension stabilizer tableau engine.
 *
 * The reference (events555/sdim) is pure Python; its hot path is
 *   Program.simulate -> Program._simulate_tableau -> apply_gate -> tableau_gates.apply_*
 *   -> ExtendedTableau.* -> tableau_optimized.*_optimized / ExtendedTableau.measure
 * (sdim/program.py:206-385, sdim/tableau/tableau_gates.py, sdim/tableau/tableau_prime.py,
 *  sdim/tableau/tableau_optimized.py).  Each entry point below says which part of that
 * path it replaces.  Plain C types only; the Python package binds it with ctypes.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative SDIM_E* code on error;
 *     sdim_last_error() then describes the error (thread-local string).
 *   - Inputs are caller-owned and copied during the call; outputs are written into
 *     caller buffers; device memory is owned by the context.
 *   - One context may be used from one host thread at a time (calls lock internally).
 */
#ifndef SDIM_HIP_H
#define SDIM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDIM_OK 0
#define SDIM_EINVAL -1     /* bad argument: gate id, qudit index, dimension, ... */
#define SDIM_EHIP -2       /* a HIP runtime call failed */
#define SDIM_ENOMEM -3     /* device memory exhausted / over budget */
#define SDIM_ENOTFOUND -4  /* unknown ticket or job */
#define SDIM_EBUSY -5      /* ticket not finished (sdim_poll) */
#define SDIM_EINTERNAL -6  /* an internal error of the library (message in sdim_last_error); host
                              memory exhaustion inside an entry point is SDIM_ENOMEM */

#define SDIM_MODE_BATCHED 0  /* persistent kernel walking device-resident gate lists */
#define SDIM_MODE_PLAIN 1    /* one kernel launch per primitive op */

typedef struct sdim_ctx sdim_ctx;

/*
 * One circuit to simulate `shots` times.  Replaces the inputs of
 * Program.simulate(shots, force_tableau=True) (sdim/program.py:206-267):
 *   ir          [n_ops][3] int64 (gate_id, qudit_index, target_index | -1), the IR triple
 *               of Program._build_ir (sdim/program.py:512-516); gate ids follow
 *               sdim/gatedata.py:42-102 (I0 X1 X_INV2 Z3 Z_INV4 H5 H_INV6 P7 P_INV8 CNOT9
 *               CNOT_INV10 CZ11 CZ_INV12 SWAP13 M14 M_X15 RESET16 N1 17).
 *   noise_channel / noise_prob  per op, read for gate 17 only ('d'=0, 'f'=1, 'p'=2; the
 *               probability that the N1 op applies a non-identity Pauli), may be NULL
 *               when the circuit has no N1 op.
 *   seed, stream  Philox4x64-10 key of every random outcome and noise draw; draw
 *               (op_index, shot) uses counter (op_index, shot, 0, 0).
 *   init        optional initial tableau in the reference layout (NULL = |0...0>):
 *               init[0..5] = x, z, ph, dx, dz, dph int64 arrays as in
 *               sdim/tableau/tableau_prime.py:24-26 (x[qudit*n + generator]).
 */
typedef struct {
  int32_t n;
  int32_t d;
  int64_t n_ops;
  const int64_t* ir;
  const int32_t* noise_channel;
  const double* noise_prob;
  int64_t shots;
  uint64_t seed;
  uint64_t stream;
  const int64_t* init[6];
} sdim_job;

/* Context on HIP device `device`.  (No reference counterpart: the reference has no device.) */
int sdim_ctx_create(int32_t device, sdim_ctx** out);
int sdim_ctx_destroy(sdim_ctx* ctx);

/* Description of the last error on this thread (replaces the reference's ValueError
 * messages, e.g. "Invalid gate value" sdim/program.py:381-382). */
const char* sdim_last_error(void);

/*
 * Tuning / engine options (no reference counterpart).  Keys:
 *   "threads"         workgroup size of the batch kernel: 0 (per launch class), 256, 512 or 1024
 *   "lds_bytes"       upper bound of the LDS budget per workgroup of the one-workgroup launch
 *                     classes (16384..163584): the LDS/HBM cut-over; gang members always take a
 *                     CU's whole LDS
 *   "shots_per_item"  shots that share one X/Z block per work item (0 = automatic: enough
 *                     items to fill the device twice, but at least min(64, n/16) shots per
 *                     item; 1 = a separate tableau per shot)
 *   "grid"            persistent grid size (0 = automatic)
 *   "allow_lds"       0 forces every tableau to stay in HBM
 *   "hbm_threads"     workgroup size of the launch class for small HBM-resident items (64..1024,
 *                     default 256)
 *   "hbm_per_cu"      LDS budget of that class is 160 KB / hbm_per_cu (default 4); at most
 *                     hbm_per_cu workgroups of it share a CU (fewer when registers limit)
 *   "serve_hbm_threads", "serve_hbm_per_cu"  the same two for the submissions of the serving
 *                     loop (defaults 512 and 2: larger workgroups, lower request latency)
 *   "gang_min_n"      circuits with at least this many qudits run on a gang of workgroups,
 *                     each owning a range of generator columns (0 = never; default 1536)
 *   "gang_chunk"      minimum bytes of a row each gang member owns (multiple of 16; default 128,
 *                     one cache line, so that no two members write the same line)
 *   "gang_threads"    workgroup size of the gang kernel: 256, 512 (default) or 1024
 *   "gang_max"        most workgroups per gang (0 = one per CU)
 *   "frame_lds_bytes" frame sampler: LDS for the frames of one workgroup's 256 shots (one
 *                     byte per qudit and shot; default 65536); larger frames live in device memory
 *   "gang_timeout"    ticks of the 100 MHz constant clock a gang member waits at a barrier before
 *                     the gang gives up and the ticket fails (default 1e9 = 10 s)
 *   "snapshots"       0 (default), 1: keep the tableau after every measurement event, 2: after every
 *                     instruction (sdim_fetch_snapshot); such submissions run without gangs
 *   "serial_max"      gate segments: a run of two or more consecutive layers of at most this many
 *                     gates (gates on disjoint qudits) becomes one serial layer that each column word
 *                     applies in program order without a barrier per layer (0 = off; default 4)
 *   "list_cap"        diagnostics: caps the kernels' LDS lists of touched rows and factor groups
 *                     (0 = their full size); an overflow drops the write and fails the ticket
 */
int sdim_set_option(sdim_ctx* ctx, const char* key, int64_t value);
int sdim_get_option(sdim_ctx* ctx, const char* key, int64_t* value);

/*
 * Simulate `n_jobs` circuits.  Replaces Program._simulate_tableau's shot loop
 * (sdim/program.py:269-365) for many programs at once.  Asynchronous: returns a
 * ticket; results are available after sdim_wait.
 */
int sdim_submit(sdim_ctx* ctx, const sdim_job* jobs, int32_t n_jobs, int32_t mode, uint64_t* ticket);
/*
 * Pauli-frame sampling of `n_jobs` circuits: the reference's default path for shots > 1
 * (Program.simulate, sdim/program.py:244-265, and simulate_frame, :45-165).  refs[j]
 * holds the reference shot's records of job j (n_meas 33-byte MEASUREMENT_DTYPE records
 * in event order, as sdim_fetch_records writes them for a one-shot, N1-free tableau run;
 * the reference's `reference_results`).  Each of the job's `shots` further shots tracks
 * a Pauli frame: x = 0 and z uniform on every qudit, conjugated by every gate; M / M_X
 * report (reference value + x_q) mod d with the reference's deterministic flag and draw
 * z_q again, RESET reports the reference record and sets x_q = 0 with a fresh z_q, N1
 * adds its (a, b) exponents.  Shots are numbered shot0, shot0 + 1, ... (records and
 * Philox counters); draws: initial z_q counter (q, shot, 1, 0), redraw after the
 * measurement at instruction k (k, shot, 1, 1), N1 (k, shot, 0, 0) as in the tableau
 * path.  Jobs start from |0...0> (init must be NULL).  Records are fetched with
 * sdim_num_measurements / sdim_fetch_records; such a ticket keeps no tableaux.
 * refs == NULL (shot0 must be 0): the library simulates each job's reference shot itself as
 * shot 0 (a tableau run with N1 as the identity, sdim/program.py:244-247, concurrently with
 * the frames on another stream) and the job's `shots` counts it: records are the full
 * [event][shot] table of Program.simulate's frame branch, and sdim_fetch_tableau(shot 0)
 * returns the reference shot's final tableau.
 */
int sdim_frame_submit(sdim_ctx* ctx, const sdim_job* jobs, int32_t n_jobs, const void* const* refs, int64_t shot0,
                      uint64_t* ticket);

/*
 * Batched simulation that allocates nothing: a variant of sdim_submit (batched mode) for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 *   out        caller-owned device buffer of out_capacity bytes.  Job j's records come first,
 *              at offset r_j (r_0 = 0, r_{j+1} = r_j + n_meas_j * shots_j rounded up to 256):
 *              n_meas_j x shots_j bytes, event-major, each value | deterministic << 7; the final
 *              tableaux follow (in the engine's packed layout).  out == NULL: only *out_bytes is
 *              set (the bytes the jobs need); no device work.
 *   out_bytes  optional; receives the bytes the jobs need.
 * The launch metadata lives in a per-GPU device array of 8 MiB (SDIM_NOALLOC_SCRATCH_BYTES)
 * held by the library; jobs whose metadata exceeds it, that need more than `out_capacity`,
 * or whose shots do not fit one work item (at most 4096, and the LDS scratch of one item)
 * are rejected with an error.  Every job runs as one work item on one workgroup (no gangs).
 * The call waits for the work (on the context's first stream) before it returns.  Calls
 * that overlap on one GPU (from different contexts) share that array and race.
 */
int sdim_simulate_noalloc(sdim_ctx* ctx, const sdim_job* jobs, int32_t n_jobs, void* out, uint64_t out_capacity,
                          uint64_t* out_bytes);
int sdim_wait(sdim_ctx* ctx, uint64_t ticket);
int sdim_poll(sdim_ctx* ctx, uint64_t ticket, int32_t* done);
int sdim_release(sdim_ctx* ctx, uint64_t ticket);

/* Number of measurement events of job `job` (M, M_X and RESET ops) - per shot. */
int sdim_num_measurements(sdim_ctx* ctx, uint64_t ticket, int32_t job, int64_t* count);

/*
 * Measurement records of job `job`: shots * M records, event-major then shot, each a
 * packed 33-byte record byte-compatible with MEASUREMENT_DTYPE (sdim/program.py:34-40):
 * (int64 qudit_index, int64 meas_round, int64 shot, bool deterministic, int64 value).
 * Replaces the MeasurementResult bookkeeping of sdim/program.py:321-332.
 */
int sdim_fetch_records(sdim_ctx* ctx, uint64_t ticket, int32_t job, void* out, int64_t capacity);

/*
 * Final tableau of (job, shot) in the reference layout: x, z, dx, dz are n*n int64
 * (row = qudit, column = generator), ph and dph n int64, all reduced as after
 * Tableau.modulo (sdim/tableau/dataclasses.py:41-47, tableau_prime.py:88-95).
 */
int sdim_fetch_tableau(sdim_ctx* ctx, uint64_t ticket, int32_t job, int64_t shot,
                       int64_t* x, int64_t* z, int64_t* ph, int64_t* dx, int64_t* dz, int64_t* dph);

/*
 * Tableau snapshot `index` of (job, shot) in the reference layout (as sdim_fetch_tableau).
 * Only for tickets submitted with the "snapshots" option: 1 keeps the tableau right after
 * every measurement event (index = event, before RESET's fix-up; record_tableau,
 * sdim/program.py:322-324), 2 after every instruction (index = instruction; the
 * reference's verbose output, sdim/program.py:313-350).
 */
int sdim_fetch_snapshot(sdim_ctx* ctx, uint64_t ticket, int32_t job, int64_t index, int64_t shot,
                        int64_t* x, int64_t* z, int64_t* ph, int64_t* dx, int64_t* dz, int64_t* dph);

/*
 * Timing and traffic of a finished ticket: device time of the simulation kernels
 * (HIP events on the launch stream) and the tableau bytes the kernels read and
 * wrote, split by where the tableau lived (HBM-resident vs LDS-resident items).
 */
int sdim_ticket_stats(sdim_ctx* ctx, uint64_t ticket, double* kernel_ms, uint64_t* bytes_hbm,
                      uint64_t* bytes_lds, int64_t* n_items, int64_t* n_lds_items);

/*
 * How a ticket was laid out on the device, by key: "items", "lds_items" (items whose
 * tableau stayed in LDS), "gang_items" (items run by gangs of workgroups), "gang_members"
 * (largest gang), "gangs" (gangs launched), "launches" (persistent launches).
 * Diagnostics only; no reference counterpart.
 */
int sdim_ticket_info(sdim_ctx* ctx, uint64_t ticket, const char* key, int64_t* value);

/*
 * Raw device counters of the last run of a ticket: [0] tableau bytes of HBM-resident
 * items, [1] of LDS-resident items, then per-phase clock ticks when the library is
 * built with SDIM_PHASE_TIMERS (zero otherwise; that build is currently switched off).
 * Returns the number of values written.
 * Diagnostics only; no reference counterpart.
 */
int sdim_ticket_counters(sdim_ctx* ctx, uint64_t ticket, uint64_t* out, int32_t count);

/*
 * Re-run the device work of a finished ticket `iters` times back to back, with every
 * input already resident in HBM (the batched path replays a captured hipGraph: reset
 * kernel + persistent kernels), and return the device time of the whole run in ms
 * (HIP events on the launch stream).  which: 0 = everything, 1 = only HBM-resident
 * items, 2 = only LDS-resident items (batched mode).  Results are recomputed
 * identically (every draw is counter-based).  No reference counterpart: used to time
 * the hot path without host-side lowering and copies.
 */
int sdim_replay(sdim_ctx* ctx, uint64_t ticket, int32_t iters, int32_t which, double* elapsed_ms);

/*
 * Overwrite every measurement record of a finished ticket (device and host copies) with
 * the byte `value` and clear its counters, so that a following sdim_replay must
 * recompute them.  Diagnostics only; no reference counterpart.
 */
int sdim_ticket_poison(sdim_ctx* ctx, uint64_t ticket, int32_t value);

/* ------------------------------------------------------------------ serving
 * Continuous batching of independent requests (one circuit each, starting from
 * |0...0>).  No reference counterpart: the reference simulates one Program at a time.
 * A dispatcher thread admits pending requests in arrival order while their estimated
 * device footprint (tableaux + records + gate lists) fits `hbm_budget_bytes`, launches
 * each admitted group as one submission on a free stream (at most `max_inflight`
 * submissions in flight, each on its own stream, so new requests overlap running
 * ones), and on completion copies the records to host memory and frees the device
 * memory.  Each request keeps its own Philox key (seed, stream): its results are the
 * same alone or in any mix.
 */
typedef struct sdim_server sdim_server;

int sdim_server_create(sdim_ctx* ctx, uint64_t hbm_budget_bytes, int32_t max_inflight, int32_t max_batch,
                       sdim_server** out);
/* Copies the request; its arrival time (for latency) is the time of this call. */
int sdim_server_submit(sdim_server* server, const sdim_job* job, uint64_t* request_id);
/* state: 0 pending, 1 on the device, 2 done, 3 failed */
int sdim_server_poll(sdim_server* server, uint64_t request_id, int32_t* state);
int sdim_server_wait(sdim_server* server, uint64_t request_id, double timeout_s);
/* Records as in sdim_fetch_records (event-major, then shot); out = NULL queries n_records.
 * A successful fetch with out != NULL releases the request. */
int sdim_server_result(sdim_server* server, uint64_t request_id, void* out, int64_t capacity,
                       int64_t* n_records, double* latency_s);
/* Completed / pending / in-flight counts, gate-shots of completed requests, time from the
 * first arrival to the last completion, and the p50 / p99 request latency (arrival to
 * records on the host). */
int sdim_server_stats(sdim_server* server, int64_t* completed, int64_t* pending, int64_t* inflight,
                      uint64_t* gate_shots, double* span_s, double* p50_s, double* p99_s);
/* Drains every pending and in-flight request, then stops the dispatcher. */
int sdim_server_destroy(sdim_server* server);

#ifdef __cplusplus
}
#endif
#endif /* SDIM_HIP_H */
