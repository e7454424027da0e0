/*
 * Serving loop of libgpu_msm.so: continuous batching of verify requests.
 *
 * No counterpart in the reference, whose callers make one blocking batch call at a time
 * (ic-integration/gpu_crypto/lib.rs:45-69, certifier loop patch:98-120).  Requests
 * (compressed signatures, -H(m) in the reference form and registered key handles, as
 * blsgpu_verify_compressed takes them) are queued on submit.  A scheduler thread owns
 * `slots` device workspaces, each with its own pair of streams.  Whenever a slot is
 * free it takes the queued requests in arrival order, up to the slot's item capacity,
 * and launches one verification step for them.  Requests that arrive while other steps
 * are on the GPU join the next free slot, so steps of different slots overlap on the
 * device.  All workspaces are allocated at creation within `hbm_budget_bytes`: the
 * item capacity of a slot is the budget divided by the slots and the bytes one item
 * needs, capped at max_step_items.  The per-key line tables are the library's resident
 * ones (blsgpu_register_key*), shared by all slots.  Results are bit-identical to
 * blsgpu_verify_compressed on the same items.
 *
 * Thread-safety: submit, poll and wait may be called from any thread.
 */
#ifndef BLSGPU_SERVE_H
#define BLSGPU_SERVE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct blsgpu_server blsgpu_server;

/* slots >= 1; max_step_items >= 1; hbm_budget_bytes > 0; max_requests_per_step = 0 for
 * no limit (1 gives the plain one-request-per-launch form); want_gt != 0 keeps the GT
 * element of every item (576 bytes each).  NULL on failure (blsgpu_last_error). */
blsgpu_server* blsgpu_server_create(int slots, int max_step_items, int64_t hbm_budget_bytes,
                                    int max_requests_per_step, int want_gt);
/* items one step of a slot can hold (after the budget), or -1 */
int blsgpu_server_step_capacity(blsgpu_server* srv);
/* Queues n items (inputs are copied).  Returns a request id >= 0, or -1 if n is out of
 * [1, step capacity], a key handle is not registered or the server is stopping. */
int64_t blsgpu_server_submit(blsgpu_server* srv, const uint8_t* sigs48, const uint64_t* neg_hms,
                             const int32_t* key_idx, int n);
/* 1: done, results (n int32, as blsgpu_verify_compressed) and gt (n x 72 u64, if the
 * server keeps them; nullable) are written, latency_ms (nullable) = completion - submit,
 * and the request is released; 0: not done yet; -1: unknown id or the step failed. */
int blsgpu_server_poll(blsgpu_server* srv, int64_t id, int32_t* results, uint64_t* gt, double* latency_ms);
/* as poll, blocking up to timeout_ms (< 0: no limit); 0 on time-out */
int blsgpu_server_wait(blsgpu_server* srv, int64_t id, int32_t* results, uint64_t* gt, double* latency_ms,
                       double timeout_ms);
/* out[0] steps launched, [1] items launched, [2] requests completed, [3] largest step
 * (items), [4] most steps in flight at once, [5] step capacity, [6] slots, [7] failed steps */
int blsgpu_server_stats(blsgpu_server* srv, double out[8]);
/* Drops queued requests, waits for the steps on the GPU, frees everything.  No other
 * call on srv may be in progress when it starts, or follow it. */
void blsgpu_server_destroy(blsgpu_server* srv);

/* Load generator for measurements: Poisson arrivals at items_per_s for `seconds`, request
 * sizes lognormal (median size_median, log-sd size_sigma, clipped to [size_min,
 * size_max]), items taken in turn from a pool of prepared items.  After the window it
 * waits up to drain_s for outstanding requests; those still open count as infinitely
 * late.  out[0] requests submitted, [1] items submitted, [2] requests completed,
 * [3] items completed, [4] seconds from first submit to last completion, [5] completed
 * items per second over that span, [6] p50, [7] p99, [8] p999, [9] max latency in ms,
 * [10] offered items per second achieved by the generator. */
int blsgpu_server_loadtest(blsgpu_server* srv, const uint8_t* pool_sigs48, const uint64_t* pool_neg_hms,
                           const int32_t* pool_key_idx, int pool_n, double items_per_s, double seconds,
                           double drain_s, double size_median, double size_sigma, int size_min, int size_max,
                           uint64_t seed, double out[11]);

#ifdef __cplusplus
}
#endif
#endif
