/*
 * params.h -- `key value` parameter-file parser (fraimondo/cudamica
 * include/params.h:17-29, src/params.c).  Same entry points and return
 * codes.  Keys are matched on the first token of a line exactly (the
 * reference tests a substring first, src/params.c:12-17, which lets e.g.
 * `afile` match inside `datafile` lines); trailing newlines are trimmed
 * instead of dropping the last character (params.c:205); the list
 * accessors always set *size, and a *size above 0 on entry is the capacity
 * of result: a longer list stores nothing and returns ERRORINVALIDPARAM
 * (the reference writes past the array; callers that pass a zeroed *size
 * keep its unbounded behaviour).  Lines starting with '#' are comments.
 */
#ifndef __PARAMS_H__
#define __PARAMS_H__

#include "error.h"
#include "config.h"

typedef struct parsedconfigs_t {
    char** configs;
    int lines;
} parsedconfigs;

#ifdef __cplusplus
extern "C" {
#endif

error parseConfig(char* filename, parsedconfigs** parsed_configs);
error freeParsedConfig(parsedconfigs* parsed_configs);
error getReal(parsedconfigs* config, char* string, real* result);
error getBool(parsedconfigs* config, char* string, natural* result);
error getInt(parsedconfigs* config, char* string, int* result);
error getLong(parsedconfigs* config, char* string, long* result);
error getNatural(parsedconfigs* config, char* string, natural* result);
error getNaturalList(parsedconfigs* config, char* string, natural* result, natural* size);
error getIntList(parsedconfigs* config, char* string, int* result, int* size);
error getLongList(parsedconfigs* config, char* string, long* result, long* size);
error getString(parsedconfigs* config, char* string, char** result);
char* getParam(char* needle, char* haystack[], int count);
int   isParam(char* needle, char* haystack[], int count);

#ifdef __cplusplus
}
#endif

#endif
