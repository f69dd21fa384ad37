/*
 * settings.h -- scalar types of the module API (reference settings.h:5-27).
 * The library is always built for complex fields: real = double,
 * field = double complex, FIELD_COMPLEX defined.
 */
#ifndef SETTINGS_H
#define SETTINGS_H

#ifndef FIELD_COMPLEX
#define FIELD_COMPLEX
#endif

#include <complex.h>

typedef double real;
typedef double complex field;

#define ABS(x) cabs(x)
#define ABS2(x) my_abs2(x)
#define CONJ(x) conj(x)
#define REAL(x) creal(x)
#define IMAG(x) cimag(x)
#define EXP(x) cexp(x)

/* |x|^2 (settings.c:4-12) */
real my_abs2(field x);

#endif
