/*
 * sn_config.c -- recursive-descent reader for the libconfig subset described
 * in sn_config.h.
 */
#include "sn_config.h"

#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const char *s;
    size_t pos, len;
    int line;
    snc_config *cfg;
    int failed;
} parser;

static void perr(parser *p, const char *msg)
{
    if (p->failed) return;
    p->failed = 1;
    snprintf(p->cfg->error, sizeof(p->cfg->error), "line %d: %s", p->line, msg);
    p->cfg->error_line = p->line;
}

static int peekc(parser *p) { return p->pos < p->len ? (unsigned char)p->s[p->pos] : -1; }
static int getc_(parser *p)
{
    if (p->pos >= p->len) return -1;
    int c = (unsigned char)p->s[p->pos++];
    if (c == '\n') p->line++;
    return c;
}

static void skip_ws(parser *p)
{
    for (;;) {
        int c = peekc(p);
        if (c < 0) return;
        if (isspace(c)) { getc_(p); continue; }
        if (c == '#') { while (peekc(p) >= 0 && peekc(p) != '\n') getc_(p); continue; }
        if (c == '/' && p->pos + 1 < p->len && p->s[p->pos + 1] == '/') {
            while (peekc(p) >= 0 && peekc(p) != '\n') getc_(p);
            continue;
        }
        if (c == '/' && p->pos + 1 < p->len && p->s[p->pos + 1] == '*') {
            getc_(p); getc_(p);
            while (peekc(p) >= 0 && !(peekc(p) == '*' && p->pos + 1 < p->len && p->s[p->pos + 1] == '/')) getc_(p);
            if (peekc(p) < 0) { perr(p, "unterminated comment"); return; }
            getc_(p); getc_(p);
            continue;
        }
        return;
    }
}

static snc_node *new_node(parser *p, snc_type t)
{
    snc_node *n = (snc_node *)calloc(1, sizeof(snc_node));
    if (!n) { perr(p, "out of memory"); return NULL; }
    n->type = t;
    n->line = p->line;
    return n;
}

static void add_kid(parser *p, snc_node *parent, snc_node *kid)
{
    snc_node **k = (snc_node **)realloc(parent->kids, sizeof(snc_node *) * (parent->nkids + 1));
    if (!k) { perr(p, "out of memory"); return; }
    parent->kids = k;
    parent->kids[parent->nkids++] = kid;
}

static void free_node(snc_node *n)
{
    if (!n) return;
    for (int i = 0; i < n->nkids; i++) free_node(n->kids[i]);
    free(n->kids);
    free(n->name);
    free(n->sval);
    free(n);
}

static snc_node *parse_value(parser *p);

static int parse_settings(parser *p, snc_node *group, int closing)
{
    for (;;) {
        skip_ws(p);
        int c = peekc(p);
        if (c < 0) {
            if (closing) { perr(p, "unexpected end of file in group"); return -1; }
            return 0;
        }
        if (closing && c == closing) { getc_(p); return 0; }
        if (!(isalpha(c) || c == '*')) { perr(p, "expected a setting name"); return -1; }
        size_t st = p->pos;
        while (peekc(p) >= 0 && (isalnum(peekc(p)) || peekc(p) == '_' || peekc(p) == '-' || peekc(p) == '*')) getc_(p);
        size_t nlen = p->pos - st;
        skip_ws(p);
        c = getc_(p);
        if (c != '=' && c != ':') { perr(p, "expected '=' or ':' after setting name"); return -1; }
        snc_node *v = parse_value(p);
        if (!v) return -1;
        v->name = (char *)malloc(nlen + 1);
        if (!v->name) { free_node(v); perr(p, "out of memory"); return -1; }
        memcpy(v->name, p->s + st, nlen);
        v->name[nlen] = 0;
        for (int i = 0; i < group->nkids; i++)
            if (group->kids[i]->name && strcmp(group->kids[i]->name, v->name) == 0) {
                char m[160];
                snprintf(m, sizeof(m), "duplicate setting '%s'", v->name);
                free_node(v);
                perr(p, m);
                return -1;
            }
        add_kid(p, group, v);
        skip_ws(p);
        c = peekc(p);
        if (c == ';' || c == ',') getc_(p);
        if (p->failed) return -1;
    }
}

static snc_node *parse_scalar(parser *p)
{
    int c = peekc(p);
    if (c == '"') {
        snc_node *n = new_node(p, SNC_STRING);
        if (!n) return NULL;
        size_t cap = 64, len = 0;
        char *buf = (char *)malloc(cap);
        if (!buf) { free_node(n); perr(p, "out of memory"); return NULL; }
        while (peekc(p) == '"') {
            getc_(p);
            for (;;) {
                int ch = getc_(p);
                if (ch < 0 || ch == '\n') { free(buf); free_node(n); perr(p, "unterminated string"); return NULL; }
                if (ch == '"') break;
                if (ch == '\\') {
                    int e = getc_(p);
                    ch = e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'f' ? '\f' : e;
                }
                if (len + 2 > cap) {
                    cap *= 2;
                    char *nb = (char *)realloc(buf, cap);
                    if (!nb) { free(buf); free_node(n); perr(p, "out of memory"); return NULL; }
                    buf = nb;
                }
                buf[len++] = (char)ch;
            }
            skip_ws(p);   /* adjacent literals concatenate */
        }
        buf[len] = 0;
        n->sval = buf;
        return n;
    }
    size_t st = p->pos;
    while (peekc(p) >= 0 && (isalnum(peekc(p)) || strchr("+-._", peekc(p)))) getc_(p);
    size_t len = p->pos - st;
    if (len == 0 || len > 63) { perr(p, "expected a value"); return NULL; }
    char tok[64];
    memcpy(tok, p->s + st, len);
    tok[len] = 0;
    if (strcasecmp(tok, "true") == 0 || strcasecmp(tok, "false") == 0) {
        snc_node *n = new_node(p, SNC_BOOL);
        if (n) n->ival = strcasecmp(tok, "true") == 0;
        return n;
    }
    char *end = NULL;
    int is_hex = (tok[0] == '0' && (tok[1] == 'x' || tok[1] == 'X')) ||
                 ((tok[0] == '-' || tok[0] == '+') && tok[1] == '0' && (tok[2] == 'x' || tok[2] == 'X'));
    int is_float = !is_hex && (strpbrk(tok, ".eE") != NULL);
    errno = 0;
    if (is_float) {
        double v = strtod(tok, &end);
        if (*end || errno) { perr(p, "malformed float"); return NULL; }
        snc_node *n = new_node(p, SNC_FLOAT);
        if (n) n->fval = v;
        return n;
    }
    long long v = strtoll(tok, &end, is_hex ? 16 : 10);
    if (*end == 'L' && end[1] == 0) end++;
    if (*end == 'L' && end[1] == 'L' && end[2] == 0) end += 2;
    if (*end || errno) { perr(p, "malformed integer"); return NULL; }
    snc_node *n = new_node(p, SNC_INT);
    if (n) n->ival = v;
    return n;
}

static snc_node *parse_value(parser *p)
{
    skip_ws(p);
    int c = peekc(p);
    if (c == '{') {
        getc_(p);
        snc_node *g = new_node(p, SNC_GROUP);
        if (!g) return NULL;
        if (parse_settings(p, g, '}')) { free_node(g); return NULL; }
        return g;
    }
    if (c == '[' || c == '(') {
        getc_(p);
        int close = (c == '[') ? ']' : ')';
        snc_node *a = new_node(p, c == '[' ? SNC_ARRAY : SNC_LIST);
        if (!a) return NULL;
        for (;;) {
            skip_ws(p);
            if (peekc(p) == close) { getc_(p); break; }
            snc_node *v = parse_value(p);
            if (!v) { free_node(a); return NULL; }
            if (a->type == SNC_ARRAY) {
                if (v->type > SNC_STRING) { free_node(v); free_node(a); perr(p, "arrays hold scalars only"); return NULL; }
                if (a->nkids && v->type != a->kids[0]->type) {
                    free_node(v); free_node(a); perr(p, "array elements must share one type"); return NULL;
                }
            }
            add_kid(p, a, v);
            skip_ws(p);
            if (peekc(p) == ',') { getc_(p); continue; }
            if (peekc(p) == close) { getc_(p); break; }
            free_node(a);
            perr(p, "expected ',' or end of array / list");
            return NULL;
        }
        return a;
    }
    if (c < 0) { perr(p, "unexpected end of file"); return NULL; }
    return parse_scalar(p);
}

int snc_read_string(snc_config *cfg, const char *text)
{
    memset(cfg, 0, sizeof(*cfg));
    parser p = {text, 0, strlen(text), 1, cfg, 0};
    snc_node *root = new_node(&p, SNC_GROUP);
    if (!root) return -1;
    if (parse_settings(&p, root, 0) || p.failed) {
        free_node(root);
        return -1;
    }
    cfg->root = root;
    return 0;
}

int snc_read_file(snc_config *cfg, const char *path)
{
    memset(cfg, 0, sizeof(*cfg));
    FILE *f = fopen(path, "rb");
    if (!f) {
        snprintf(cfg->error, sizeof(cfg->error), "cannot open %s", path);
        return -1;
    }
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc(n + 1);
    if (!buf) { fclose(f); snprintf(cfg->error, sizeof(cfg->error), "out of memory"); return -1; }
    size_t got = fread(buf, 1, n, f);
    fclose(f);
    buf[got] = 0;
    int r = snc_read_string(cfg, buf);
    free(buf);
    return r;
}

void snc_destroy(snc_config *cfg)
{
    free_node(cfg->root);
    cfg->root = NULL;
}

const snc_node *snc_lookup(const snc_config *cfg, const char *path)
{
    const snc_node *cur = cfg->root;
    const char *s = path;
    while (cur && *s) {
        const char *dot = strchr(s, '.');
        size_t len = dot ? (size_t)(dot - s) : strlen(s);
        const snc_node *next = NULL;
        if (cur->type == SNC_GROUP) {
            for (int i = 0; i < cur->nkids; i++)
                if (cur->kids[i]->name && strlen(cur->kids[i]->name) == len &&
                    strncmp(cur->kids[i]->name, s, len) == 0) { next = cur->kids[i]; break; }
        } else if ((cur->type == SNC_LIST || cur->type == SNC_ARRAY) && s[0] == '[') {
            int idx = atoi(s + 1);
            if (idx >= 0 && idx < cur->nkids) next = cur->kids[idx];
        }
        cur = next;
        s = dot ? dot + 1 : s + len;
    }
    return cur;
}

int snc_lookup_int64(const snc_config *cfg, const char *path, int64_t *out)
{
    const snc_node *n = snc_lookup(cfg, path);
    if (!n || n->type != SNC_INT) return 0;
    *out = n->ival;
    return 1;
}

int snc_lookup_int(const snc_config *cfg, const char *path, int *out)
{
    int64_t v;
    if (!snc_lookup_int64(cfg, path, &v) || v < -2147483648LL || v > 2147483647LL) return 0;
    *out = (int)v;
    return 1;
}

int snc_lookup_float(const snc_config *cfg, const char *path, double *out)
{
    const snc_node *n = snc_lookup(cfg, path);
    if (!n || n->type != SNC_FLOAT) return 0;
    *out = n->fval;
    return 1;
}

int snc_lookup_bool(const snc_config *cfg, const char *path, int *out)
{
    const snc_node *n = snc_lookup(cfg, path);
    if (!n || n->type != SNC_BOOL) return 0;
    *out = (int)n->ival;
    return 1;
}

int snc_lookup_string(const snc_config *cfg, const char *path, const char **out)
{
    const snc_node *n = snc_lookup(cfg, path);
    if (!n || n->type != SNC_STRING) return 0;
    *out = n->sval;
    return 1;
}

int snc_length(const snc_node *n)
{
    if (!n || (n->type != SNC_ARRAY && n->type != SNC_LIST && n->type != SNC_GROUP)) return 0;
    return n->nkids;
}

double snc_float_elem(const snc_node *n, int i)
{
    if (!n || i < 0 || i >= n->nkids) return 0.0;
    const snc_node *e = n->kids[i];
    return e->type == SNC_FLOAT ? e->fval : 0.0;
}
