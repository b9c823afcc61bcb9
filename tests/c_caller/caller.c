/* A plain C99 caller of the library's C ABI: includes every public header,
 * links against libkinet_crypto_hip.so and exercises the argument checks that
 * need no device.  Built and run by tests/test_capi.py. */
#include <stdio.h>
#include <string.h>

#include "kinet/crypto/bls.h"
#include "kinet/crypto/bn254.h"
#include "kinet/crypto/hip_runtime.h"
#include "kinet/crypto/kzg.h"
#include "kinet/crypto/kzg_serve.h"
#include "kinet/crypto/ntt.h"
#include "kinet/gpukit/batch_inversion.h"
#include "kinet/gpukit/gpukit.h"
#include "kinet/gpukit/multi_pippenger.h"

static int fails = 0;
#define CHECK(c)                                              \
    do {                                                      \
        if (!(c)) {                                           \
            fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
            ++fails;                                          \
        }                                                     \
    } while (0)

int main(void) {
    uint8_t out[144];
    uint8_t buf[160];
    memset(buf, 0, sizeof buf);
    CHECK(gpukit_multi_pippenger_hip(GPUKIT_CURVE_BLS12_381_G1, buf, buf, 1, NULL) == GPUKIT_ERR_NULL_ARG);
    CHECK(gpukit_multi_pippenger_hip(4, buf, buf, 1, out) == GPUKIT_ERR_BACKEND); /* no such curve */
    CHECK(gpukit_multi_pippenger_hip(GPUKIT_CURVE_BANDERWAGON, NULL, NULL, 0, out) == GPUKIT_OK
          && out[63] == 1 && out[0] == 0); /* the identity (0, 1) */
    CHECK(gpukit_multi_pippenger_hip(GPUKIT_CURVE_BN254_G1, NULL, NULL, 0, out) == GPUKIT_OK);
    CHECK(ntt_fr_forward(KINET_NTT_FIELD_BLS12_381_FR, NULL, 4) == -1);
    CHECK(ntt_fr_forward(KINET_NTT_FIELD_BN254_FR, buf, 3) == -1); /* not a power of two */
    CHECK(kzg_srs_create(2, NULL, 0, NULL) == -1);
    CHECK(bls12_381_g1_msm(out, out + 64, buf, 159) == 0); /* not a whole entry */
    CHECK(kinet_hip_points_bytes(2, 10) == 1280 && kinet_hip_points_bytes(1, 10) == 800);
    CHECK(kinet_hip_msm_prepared_noalloc(2, buf, buf, buf, ((size_t)1 << 16) + 1, out, NULL) == GPUKIT_ERR_BAD_SIZE);
    CHECK(kinet_bn254_hip_fp_mul(buf, buf, out, 0) == 0);
    CHECK(kzg_blob_to_commit(NULL, out) == -1);
    CHECK(kzg_load_trusted_setup_file(NULL) == -1);
    CHECK(gpukit_batch_inv_bls12_381_fp_hip(buf, out, 0) == GPUKIT_OK);
    CHECK(gpukit_batch_inv_bn254_fp_hip(NULL, out, 1) == GPUKIT_ERR_NULL_ARG);
    CHECK(kinet_hip_batch_inv_fp(1, buf, buf + 16, 1, NULL, NULL) == GPUKIT_ERR_NULL_ARG);
    CHECK(kinet_hip_batch_inv_fp(1, buf, buf + 16, 1, (uint32_t*)out, buf) == GPUKIT_ERR_BAD_SIZE); /* overlap */
    printf("device available: %d\n", kinet_hip_available());
    printf("%s\n", fails ? "C caller: FAIL" : "C caller: ok");
    return fails != 0;
}
