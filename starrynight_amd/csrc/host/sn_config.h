/*
 * sn_config.h -- reader for starrynight.cfg (a subset of the libconfig
 * grammar, written for this project; libconfig itself is not used).
 *
 * Grammar covered: comments (#, //, block comments), settings `name = value`
 * or `name : value` with optional `;` / `,`, integers (decimal, 0x hex, L
 * suffix), floats, booleans (true / false, any case), strings (adjacent
 * literals concatenate, \" \\ \n \t escapes), arrays [ ... ], lists ( ... ) and
 * groups { ... }.  Lookups take dotted paths ("Efield.x").
 *
 * Typing follows libconfig's default (no auto-conversion), which the
 * reference relies on (src/starrynight-config.c:125-179, and the cfg note
 * "MCMoves ... Must be floating point!"): an int lookup of a float setting, or
 * a float lookup of an int setting, fails and leaves the caller's default.
 */
#ifndef SN_CONFIG_H
#define SN_CONFIG_H

#include <stdint.h>

typedef enum { SNC_INT, SNC_FLOAT, SNC_BOOL, SNC_STRING, SNC_ARRAY, SNC_LIST, SNC_GROUP } snc_type;

typedef struct snc_node {
    char *name;             /* NULL for array / list elements */
    snc_type type;
    int64_t ival;
    double fval;
    char *sval;
    struct snc_node **kids; /* arrays, lists, groups */
    int nkids;
    int line;
} snc_node;

typedef struct {
    snc_node *root;         /* a group */
    char error[256];
    int error_line;
} snc_config;

/* Parse a file / a string.  Return 0 on success, -1 on error (message in cfg->error). */
int snc_read_file(snc_config *cfg, const char *path);
int snc_read_string(snc_config *cfg, const char *text);
void snc_destroy(snc_config *cfg);

const snc_node *snc_lookup(const snc_config *cfg, const char *path);
/* 1 when found with the right type (value written), 0 otherwise. */
int snc_lookup_int(const snc_config *cfg, const char *path, int *out);
int snc_lookup_int64(const snc_config *cfg, const char *path, int64_t *out);
int snc_lookup_float(const snc_config *cfg, const char *path, double *out);
int snc_lookup_bool(const snc_config *cfg, const char *path, int *out);
int snc_lookup_string(const snc_config *cfg, const char *path, const char **out);
/* Number of elements of an array / list (0 if absent). */
int snc_length(const snc_node *n);
/* Element i as float: floats only (ints give 0.0, as libconfig's
 * config_setting_get_float_elem does without auto-conversion). */
double snc_float_elem(const snc_node *n, int i);

#endif
