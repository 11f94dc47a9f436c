/*
 * error.h -- error codes of the AMICA host interface.
 * The codes and their values are those of the reference
 * (fraimondo/cudamica include/error.h:111-118).  The vendor-library check
 * macros of the reference (cuBLAS, cuRAND, MAGMA, MKL) have no counterpart:
 * HIP failures inside the engine are returned as AMICA_E_* codes and turned
 * into a message and exit(1) by C_CHECK_RETURN, as the reference does.
 */
#ifndef ERROR_H_
#define ERROR_H_

#include <stdio.h>
#include <stdlib.h>
#include "tipos.h"

#define SUCCESS 0
#define ERRORNODEVICEMEM   -1   /* No more memory on device */
#define ERRORNOPARAM       -2   /* Parameter error */
#define ERRORINVALIDPARAM  -3   /* Parameter is invalid */
#define ERRORINVALIDCONFIG -4   /* Config file is invalid */
#define ERRORNOFILE        -5   /* Error opening file */
#define ERRORIO            -6   /* Input/Output error */
#define ERRORNOMEM         -7   /* No more memory on host */

#ifdef __cplusplus
extern "C" {
#endif
/* Message for an error code (the reference's table has 6 of the 8 codes,
 * error.cu:5-16; this one has all of them). */
const char* cGetErrorString(natural value);
#ifdef __cplusplus
}
#endif

/* Prints the failing call and exits, like the reference's C_CHECK_RETURN. */
#define C_CHECK_RETURN(value) {                                              \
    error _m_cerr = (error)(value);                                          \
    if (_m_cerr != SUCCESS) {                                                \
        fprintf(stderr, "ERROR::%s at line %d in file %s\n",                 \
                cGetErrorString((natural)_m_cerr), __LINE__, __FILE__);      \
        exit(1);                                                             \
    } }

#endif /* ERROR_H_ */
