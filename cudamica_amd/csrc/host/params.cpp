/*
 * params.cpp -- `key value` parameter-file parser with the entry points of
 * fraimondo/cudamica src/params.c:7-297.
 *
 * Same contract: parseConfig reads every line of the file (lines starting
 * with '#' are comments), the get* accessors return SUCCESS, ERRORNOPARAM
 * (key absent) or ERRORINVALIDPARAM (key present without a usable value).
 * Differences, all resolving reference defects: the key is the first
 * whitespace-separated token of a line, matched exactly (the reference
 * pre-filters with strstr and tokenises the buffer destructively, so
 * results could depend on call order); tokens are split on blanks and
 * tabs; string values lose their line terminator instead of always their
 * last character (params.c:205 assumes a trailing newline); the list
 * accessors always report the count (getIntList never set it).
 */
#include "../../../include/params.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

const char* kSep = " \t\r\n";

/* Tokens of the first line whose first token is `key`; empty if none. */
bool find_key(parsedconfigs* cfg, const char* key, std::vector<std::string>* values) {
    if (!cfg || !key) return false;
    for (int i = 0; i < cfg->lines; ++i) {
        const char* line = cfg->configs[i];
        if (!line) continue;
        std::string s(line);
        std::vector<std::string> tok;
        size_t p = s.find_first_not_of(kSep);
        while (p != std::string::npos) {
            size_t q = s.find_first_of(kSep, p);
            tok.push_back(s.substr(p, q == std::string::npos ? std::string::npos : q - p));
            p = (q == std::string::npos) ? q : s.find_first_not_of(kSep, q);
        }
        if (!tok.empty() && tok[0] == key) {
            values->assign(tok.begin() + 1, tok.end());
            return true;
        }
    }
    return false;
}

}  // namespace

extern "C" {

error getReal(parsedconfigs* config, char* string, real* result) {
    std::vector<std::string> v;
    if (!find_key(config, string, &v)) return ERRORNOPARAM;
    if (v.empty()) return ERRORINVALIDPARAM;
    *result = (real)atof(v[0].c_str());
    return SUCCESS;
}

error getBool(parsedconfigs* config, char* string, natural* result) {
    std::vector<std::string> v;
    if (!find_key(config, string, &v)) return ERRORNOPARAM;
    if (v.empty()) return ERRORINVALIDPARAM;
    if (v[0] == "on") *result = 1;
    else if (v[0] == "off") *result = 0;
    else if (v[0] == "none") *result = 2;
    else return ERRORINVALIDPARAM;
    return SUCCESS;
}

error getInt(parsedconfigs* config, char* string, int* result) {
    std::vector<std::string> v;
    if (!find_key(config, string, &v)) return ERRORNOPARAM;
    if (v.empty()) return ERRORINVALIDPARAM;
    *result = atoi(v[0].c_str());
    return SUCCESS;
}

error getLong(parsedconfigs* config, char* string, long* result) {
    std::vector<std::string> v;
    if (!find_key(config, string, &v)) return ERRORNOPARAM;
    if (v.empty()) return ERRORINVALIDPARAM;
    *result = atol(v[0].c_str());
    return SUCCESS;
}

error getNatural(parsedconfigs* config, char* string, natural* result) {
    return getLong(config, string, (long*)result);
}

/* *size > 0 on entry: the capacity of result */
error getIntList(parsedconfigs* config, char* string, int* result, int* size) {
    std::vector<std::string> v;
    if (!find_key(config, string, &v)) return ERRORNOPARAM;
    if (size && *size > 0 && (long)v.size() > (long)*size) return ERRORINVALIDPARAM;
    int n = 0;
    for (const auto& s : v) result[n++] = atoi(s.c_str());
    if (size) *size = n;
    return SUCCESS;
}

error getLongList(parsedconfigs* config, char* string, long* result, long* size) {
    std::vector<std::string> v;
    if (!find_key(config, string, &v)) return ERRORNOPARAM;
    if (size && *size > 0 && (long)v.size() > *size) return ERRORINVALIDPARAM;
    long n = 0;
    for (const auto& s : v) result[n++] = atol(s.c_str());
    if (size) *size = n;
    return SUCCESS;
}

error getNaturalList(parsedconfigs* config, char* string, natural* result, natural* size) {
    return getLongList(config, string, (long*)result, (long*)size);
}

error getString(parsedconfigs* config, char* string, char** result) {
    std::vector<std::string> v;
    if (!find_key(config, string, &v)) return ERRORNOPARAM;
    if (v.empty()) return ERRORINVALIDPARAM;
    *result = (char*)malloc(v[0].size() + 1);
    std::memcpy(*result, v[0].c_str(), v[0].size() + 1);
    return SUCCESS;
}

char* getParam(char* needle, char* haystack[], int count) {
    for (int i = 0; i < count; i++)
        if (strcmp(needle, haystack[i]) == 0 && i < count - 1) return haystack[i + 1];
    return 0;
}

int isParam(char* needle, char* haystack[], int count) {
    for (int i = 0; i < count; i++)
        if (strcmp(needle, haystack[i]) == 0) return 1;
    return 0;
}

error parseConfig(char* filename, parsedconfigs** parsed_configs) {
    fprintf(stdout, "====================================\n");
    fprintf(stdout, " Opening config file %s\n", filename);
    fprintf(stdout, "====================================\n");
    FILE* f = fopen(filename, "r");
    if (!f) {
        *parsed_configs = NULL;
        return ERRORINVALIDCONFIG;
    }
    std::vector<std::string> lines;
    std::string cur;
    int ch;
    while ((ch = fgetc(f)) != EOF) {
        cur.push_back((char)ch);
        if (ch == '\n') { lines.push_back(cur); cur.clear(); }
    }
    if (!cur.empty()) lines.push_back(cur);
    fclose(f);
    parsedconfigs* r = (parsedconfigs*)malloc(sizeof(parsedconfigs));
    r->lines = (int)lines.size();
    r->configs = (char**)malloc(sizeof(char*) * (lines.empty() ? 1 : lines.size()));
    for (size_t i = 0; i < lines.size(); ++i) {
        if (!lines[i].empty() && lines[i][0] == '#') {
            r->configs[i] = NULL;
        } else {
            r->configs[i] = (char*)malloc(lines[i].size() + 1);
            std::memcpy(r->configs[i], lines[i].c_str(), lines[i].size() + 1);
        }
    }
    *parsed_configs = r;
    fprintf(stdout, "Done!\n");
    fprintf(stdout, "====================================\n\n");
    return SUCCESS;
}

error freeParsedConfig(parsedconfigs* config) {
    if (!config) return SUCCESS;
    for (int i = 0; i < config->lines; i++) free(config->configs[i]);
    free(config->configs);
    free(config);
    return SUCCESS;
}

}  // extern "C"
