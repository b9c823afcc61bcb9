// Driver of the host check under AddressSanitizer and UndefinedBehaviorSanitizer
// (make -C tests/hostcheck sanitize): exercises the host build of the device
// arithmetic (field.h, curve.h, curve_quad.h, glv.h) and the serving scheduler
// (serve_core.h) through the same entries the CPU tests call, checking
// identities that need no reference values.  Any sanitizer report, or a
// failed identity, makes it exit non-zero.  Test infrastructure only.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

extern "C" {
void hc_field_op(int field, int op, const uint8_t* a, const uint8_t* b, uint8_t* out);
void hc_point_op(int curve, int op, const uint8_t* a, const uint8_t* b, uint8_t* out);
void hc_glv_split_bn(const uint8_t* s32, uint8_t* out);
void hc_glv_split_bls(const uint8_t* s32, uint8_t* out);
int hc_quad_op(int curve, int op, int lazy, const uint8_t* a, const uint8_t* b, const uint8_t* za, const uint8_t* zb,
               const int* ka, const int* kb, uint8_t* out);
int hc_quad_tree(int curve, int lazy, const uint8_t* pts, const uint8_t* zs, int n, uint8_t* out);
int hc_quad_dbl_chain(int curve, const uint8_t* a, const uint8_t* za, const int* ka, int count, uint8_t* out);
int hc_serve_scenario(uint32_t lanes, uint32_t max_batch, uint64_t max_batch_elems, uint64_t budget, size_t nreq,
                      const int* curve, const uint32_t* n, const int* bad, const int* launch_fail,
                      const int* device_fail, int latency_us, int* status, uint64_t* commit8, int* batch,
                      uint64_t* stats_out);
}

static int failures = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                 \
        }                                                               \
    } while (0)

int main() {
    std::mt19937_64 rng(20261204);
    // field identities on small random values (below every modulus)
    const int nbytes[5] = {48, 32, 32, 32, 32};
    for (int f = 0; f < 5; ++f) {
        for (int t = 0; t < 200; ++t) {
            uint8_t a[48] = {0}, b[48] = {0}, s[48], d[48], ab[48], ba[48], inv[48], one[48];
            for (int k = 0; k < nbytes[f] - 2; ++k) {
                a[k] = (uint8_t)rng();
                b[k] = (uint8_t)rng();
            }
            a[0] |= 1;
            hc_field_op(f, 0, a, b, s);
            hc_field_op(f, 1, s, b, d);
            CHECK(std::memcmp(d, a, nbytes[f]) == 0);  // (a + b) - b = a
            hc_field_op(f, 2, a, b, ab);
            hc_field_op(f, 2, b, a, ba);
            CHECK(std::memcmp(ab, ba, nbytes[f]) == 0);
            hc_field_op(f, 3, a, b, inv);
            hc_field_op(f, 2, a, inv, one);
            CHECK(one[0] == 1);
            for (int k = 1; k < nbytes[f]; ++k) CHECK(one[k] == 0);
        }
    }
    // points: BN254 generator (1, 2); 2G by doubling, by addition, by the quad operations
    uint8_t g[64] = {0}, g2[64], g2b[64], q[64], z[32] = {7}, z2[32] = {11};
    g[31] = 1;
    g[63] = 2;
    const int k0[4] = {0, 0, 0, 0}, kmax[4] = {7, 3, 1, 1};
    hc_point_op(1, 2, g, g, g2);
    hc_point_op(1, 0, g, g, g2b);
    CHECK(std::memcmp(g2, g2b, 64) == 0);
    for (int lazy = 0; lazy < 2; ++lazy) {
        CHECK(hc_quad_op(1, 0, lazy, g, g, z, z2, k0, lazy ? kmax : k0, q) == 0);
        CHECK(std::memcmp(q, g2, 64) == 0);
        CHECK(hc_quad_op(1, 1, lazy, g, g, z, z, lazy ? kmax : k0, k0, q) == 0);
        CHECK(std::memcmp(q, g2, 64) == 0);
    }
    uint8_t pts[8 * 64], zs[8 * 32] = {0};
    for (int i = 0; i < 8; ++i) {
        std::memcpy(pts + 64 * i, g, 64);
        zs[32 * i] = (uint8_t)(3 + i);
    }
    uint8_t g8[64], chain[64];
    CHECK(hc_quad_tree(1, 1, pts, zs, 8, g8) == 0);
    CHECK(hc_quad_dbl_chain(1, g, z, kmax, 3, chain) == 0);
    CHECK(std::memcmp(g8, chain, 64) == 0);  // 8 G both ways
    // GLV splits of random scalars (magnitudes below 2^127)
    for (int t = 0; t < 2000; ++t) {
        uint8_t s[32], out[34];
        for (auto& x : s) x = (uint8_t)rng();
        hc_glv_split_bn(s, out);
        CHECK((out[15] & 0x80) == 0 && (out[31] & 0x80) == 0);
        hc_glv_split_bls(s, out);
        CHECK((out[15] & 0x80) == 0 && (out[31] & 0x80) == 0);
    }
    // the serving scheduler with a bad request, a launch failure and a device failure
    const int m = 30;
    int curve[m], bad[m], status[m], batch[m];
    uint32_t n[m];
    uint64_t c8[m], st[9];
    for (int i = 0; i < m; ++i) {
        curve[i] = 1 + i % 2;
        n[i] = 1u << (4 + i % 5);
        bad[i] = i == 7;
    }
    const int lf[2] = {2, -1}, df[2] = {4, -1};
    hc_serve_scenario(2, 4, 1u << 20, 1u << 16, m, curve, n, bad, lf, df, 100, status, c8, batch, st);
    CHECK(st[0] == (uint64_t)m && st[1] + st[2] == (uint64_t)m);
    for (int i = 0; i < m; ++i) CHECK(status[i] != 0 || c8[i] == (uint64_t)(i + 1));
    CHECK(status[7] != 0);
    std::printf("sancheck: %s (%d failures)\n", failures ? "FAIL" : "ok", failures);
    return failures ? 1 : 0;
}
