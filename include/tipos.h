/*
 * tipos.h -- scalar types of the AMICA host interface.
 * Same names and widths as the reference build (fraimondo/cudamica
 * include/tipos.h:11-26 with USELONG, which every reference build enables,
 * configure.ac:145-148): natural/integer are 64-bit, real is double.
 */
#ifndef TIPOS_H_
#define TIPOS_H_

typedef unsigned long natural;
typedef long integer;
typedef int error;
typedef int boolean;
typedef double real;

#endif /* TIPOS_H_ */
