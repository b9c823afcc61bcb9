/* Internal declarations of the CPU oracle (test infrastructure only). */
#ifndef KINET_ORC_INTERNAL_H
#define KINET_ORC_INTERNAL_H

#include "oracle.h"

typedef uint64_t u64;
typedef unsigned __int128 u128;

typedef struct {
    int n;        /* u64 limbs */
    int bits;     /* bit length of p */
    u64 p[6];
    u64 r1[6];    /* R mod p   (Montgomery one) */
    u64 r2[6];    /* R^2 mod p */
    u64 pinv;     /* -p^-1 mod 2^64 */
} orc_field;

const orc_field* orc_get_field(int id);

u64  mp_add(u64* r, const u64* a, const u64* b, int n);
u64  mp_sub(u64* r, const u64* a, const u64* b, int n);
int  mp_cmp(const u64* a, const u64* b, int n);
int  mp_is_zero(const u64* a, int n);

void f_add(const orc_field* F, u64* r, const u64* a, const u64* b);
void f_sub(const orc_field* F, u64* r, const u64* a, const u64* b);
void f_neg(const orc_field* F, u64* r, const u64* a);
void f_mul(const orc_field* F, u64* r, const u64* a, const u64* b);
void f_to_mont(const orc_field* F, u64* r, const u64* a);
void f_from_mont(const orc_field* F, u64* r, const u64* a);
void f_pow(const orc_field* F, u64* r, const u64* a, const u64* e, int ne);
void f_inv(const orc_field* F, u64* r, const u64* a);
int  f_sqrt(const orc_field* F, u64* r, const u64* a);

void le_to_limbs(u64* r, const uint8_t* b, int nbytes);
void limbs_to_le(uint8_t* b, const u64* a, int nbytes);
void be_to_limbs(u64* r, const uint8_t* b, int nbytes);
void limbs_to_be(uint8_t* b, const u64* a, int nbytes);

/* Jacobian point, coordinates in Montgomery form. */
typedef struct {
    u64 X[6], Y[6], Z[6];
    int inf;
} orc_jac;

typedef struct {
    const orc_field* F;   /* base field */
    const orc_field* R;   /* scalar field */
    int fp_bytes;         /* 32 or 48 */
    u64 b_mont[6];        /* curve constant b in Montgomery form */
} orc_curve;

const orc_curve* orc_get_curve(int curve);

void g_zero(orc_jac* r);
void g_dbl(const orc_curve* C, orc_jac* r, const orc_jac* p);
void g_add(const orc_curve* C, orc_jac* r, const orc_jac* a, const orc_jac* b);
void g_neg(const orc_curve* C, orc_jac* r, const orc_jac* p);
void g_mul(const orc_curve* C, orc_jac* r, const orc_jac* p, const u64* k4);
void g_to_affine(const orc_curve* C, u64* x, u64* y, int* inf, const orc_jac* p);
void g_read_wire(const orc_curve* C, orc_jac* r, const uint8_t* wire);
void g_write_wire(const orc_curve* C, uint8_t* wire, const orc_jac* p);
void g_generator(const orc_curve* C, orc_jac* r, int curve);

#endif
