/*
 * KZG commitment serving loop on one MI355X: requests arrive one at a time
 * (any thread), a scheduler thread admits them against an HBM budget and
 * batches them continuously onto a few HIP streams ("lanes"), each batch one
 * launch sequence (inverse NTT per size, batched MSM over the resident SRS).
 *
 * No counterpart in the reference (its KZG entry, kzg_blob_to_commit in
 * c-abi/crypto.h:238, is one synchronous call per blob).  Commitments are the
 * bytes kzg_commit (kzg.h) returns for the same evaluations: a request's
 * result does not depend on what it was batched with.
 *
 * Returns CRYPTO_OK (0) or CRYPTO_ERR_* (c-abi/crypto.h:60-66).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "kzg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kzg_server kzg_server;

typedef struct kzg_server_config {
    uint32_t lanes;           /* concurrent batches (HIP streams); 0 = 3 */
    uint32_t max_batch;       /* requests per batch; 0 = 128 */
    uint64_t max_batch_elems; /* evaluations per batch; 0 = the largest power of two within an SRS, at least 2^20 */
    uint64_t hbm_budget;      /* device bytes of admitted, unfinished batches; 0 = 8 GiB */
} kzg_server_config;

typedef struct kzg_server_stats {
    uint64_t submitted, completed, failed;
    uint64_t batches;
    uint64_t max_batch_requests;  /* largest batch seen */
    uint64_t peak_inflight_bytes; /* largest admitted footprint */
    uint64_t budget_waits;        /* times a ready batch waited for the budget */
    uint64_t rerun_batches;       /* batches re-run one request at a time after an error flag */
    uint64_t launch_failures;     /* batches whose staging or launch failed (their requests fail) */
} kzg_server_stats;

/* srs_bn254 / srs_bls12_381: resident SRS handles (either may be NULL if that
 * curve is not served).  cfg may be NULL (defaults). */
int kzg_server_create(const kzg_srs* srs_bn254, const kzg_srs* srs_bls12_381, const kzg_server_config* cfg,
                      kzg_server** out);
/* Finishes the admitted work, stops the scheduler and frees everything. */
void kzg_server_destroy(kzg_server* server);

/* Queues a commitment to the polynomial with evaluations evals (n x 32-byte
 * LE canonical Fr, natural order, n a power of two <= the curve's SRS size).
 * The caller keeps evals valid and unchanged until kzg_server_wait returns
 * for the ticket.  Returns a ticket >= 0, or a negative CRYPTO_ERR_*. */
int64_t kzg_server_submit(kzg_server* server, uint32_t curve, const uint8_t* evals, size_t n);

/* Blocks until the ticket completes; writes the compressed commitment (48 or
 * 32 bytes) and, if not NULL, the submit and completion times in nanoseconds
 * since the server started.  CRYPTO_ERR_INPUT if an evaluation was >= r. */
int kzg_server_wait(kzg_server* server, int64_t ticket, uint8_t* commit_out, uint64_t* t_submit_ns,
                    uint64_t* t_done_ns);

int kzg_server_get_stats(const kzg_server* server, kzg_server_stats* out);

#ifdef __cplusplus
} /* extern "C" */
#endif
