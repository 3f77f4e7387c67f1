/*
 * Minimal declarations of GMP's documented public API (GMP 6.x, 64-bit limbs),
 * written by hand because this image ships libgmp.so.10 without gmp.h.
 * Only the functions the oracle uses are declared.  The mpz struct layout and
 * the __gmpz_/__gmpn_ symbol names are GMP's stable public ABI.
 *
 * TEST INFRASTRUCTURE ONLY: used by the CPU oracle under oracle/.
 */
#ifndef GPUPI_ORACLE_GMP_DECL_H
#define GPUPI_ORACLE_GMP_DECL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned long mp_limb_t;
typedef long mp_size_t;
typedef unsigned long mp_bitcnt_t;

typedef struct {
    int _mp_alloc;
    int _mp_size;
    mp_limb_t *_mp_d;
} gmp_mpz_struct;
typedef gmp_mpz_struct mpz_t[1];
typedef gmp_mpz_struct *mpz_ptr;
typedef const gmp_mpz_struct *mpz_srcptr;

void __gmpz_init(mpz_ptr);
void __gmpz_clear(mpz_ptr);
void __gmpz_set_ui(mpz_ptr, unsigned long);
void __gmpz_mul_2exp(mpz_ptr, mpz_srcptr, mp_bitcnt_t);
void __gmpz_fdiv_qr(mpz_ptr, mpz_ptr, mpz_srcptr, mpz_srcptr);
void __gmpz_import(mpz_ptr, size_t, int, size_t, int, size_t, const void *);
void *__gmpz_export(void *, size_t *, int, size_t, int, size_t, mpz_srcptr);
void __gmpz_ui_pow_ui(mpz_ptr, unsigned long, unsigned long);
size_t __gmpz_sizeinbase(mpz_srcptr, int);

mp_limb_t __gmpn_mul(mp_limb_t *, const mp_limb_t *, mp_size_t, const mp_limb_t *, mp_size_t);
void __gmpn_sqr(mp_limb_t *, const mp_limb_t *, mp_size_t);
mp_limb_t __gmpn_add(mp_limb_t *, const mp_limb_t *, mp_size_t, const mp_limb_t *, mp_size_t);
mp_limb_t __gmpn_sub(mp_limb_t *, const mp_limb_t *, mp_size_t, const mp_limb_t *, mp_size_t);
mp_limb_t __gmpn_mul_1(mp_limb_t *, const mp_limb_t *, mp_size_t, mp_limb_t);
mp_limb_t __gmpn_lshift(mp_limb_t *, const mp_limb_t *, mp_size_t, unsigned int);
mp_limb_t __gmpn_rshift(mp_limb_t *, const mp_limb_t *, mp_size_t, unsigned int);
void __gmpn_divexact_1(mp_limb_t *, const mp_limb_t *, mp_size_t, mp_limb_t);
mp_limb_t __gmpn_mod_1(const mp_limb_t *, mp_size_t, mp_limb_t);

#define mpz_init __gmpz_init
#define mpz_clear __gmpz_clear
#define mpz_set_ui __gmpz_set_ui
#define mpz_mul_2exp __gmpz_mul_2exp
#define mpz_fdiv_qr __gmpz_fdiv_qr
#define mpz_import __gmpz_import
#define mpz_export __gmpz_export
#define mpz_ui_pow_ui __gmpz_ui_pow_ui
#define mpz_sizeinbase __gmpz_sizeinbase
#define mpn_mul __gmpn_mul
#define mpn_sqr __gmpn_sqr
#define mpn_add __gmpn_add
#define mpn_sub __gmpn_sub
#define mpn_mul_1 __gmpn_mul_1
#define mpn_lshift __gmpn_lshift
#define mpn_rshift __gmpn_rshift
#define mpn_divexact_1 __gmpn_divexact_1
#define mpn_mod_1 __gmpn_mod_1

#ifdef __cplusplus
}
#endif

#endif
