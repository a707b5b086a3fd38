/*
 * mmc_json.c -- minimal recursive-descent JSON reader (see mmc_json.h).
 */
#include "mmc_json.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const char* s;
    const char* err;
    int depth;
} mj_cur;

static void skip_ws(mj_cur* c) {
    for (;;) {
        while (isspace((unsigned char)*c->s)) c->s++;
        if (c->s[0] == '/' && c->s[1] == '/') { /* tolerate // comments */
            while (*c->s && *c->s != '\n') c->s++;
            continue;
        }
        break;
    }
}

static mj_node* node_new(int type) {
    mj_node* n = (mj_node*)calloc(1, sizeof(mj_node));
    if (n) n->type = type;
    return n;
}

static char* parse_string_raw(mj_cur* c) {
    if (*c->s != '"') {
        c->err = "expected a string";
        return NULL;
    }
    c->s++;
    size_t cap = 32, len = 0;
    char* out = (char*)malloc(cap);
    while (*c->s && *c->s != '"') {
        char ch = *c->s++;
        if (ch == '\\') {
            char e = *c->s++;
            switch (e) {
                case 'n': ch = '\n'; break;
                case 't': ch = '\t'; break;
                case 'r': ch = '\r'; break;
                case 'b': ch = '\b'; break;
                case 'f': ch = '\f'; break;
                case 'u': { /* keep ASCII code points, replace others */
                    unsigned v = 0;
                    for (int k = 0; k < 4 && isxdigit((unsigned char)*c->s); k++, c->s++)
                        v = v * 16 + (unsigned)(isdigit((unsigned char)*c->s) ? *c->s - '0' : (tolower(*c->s) - 'a' + 10));
                    ch = v < 128 ? (char)v : '?';
                    break;
                }
                case 0: c->err = "unterminated string"; free(out); return NULL;
                default: ch = e; break;
            }
        }
        if (len + 2 > cap) {
            cap *= 2;
            out = (char*)realloc(out, cap);
        }
        out[len++] = ch;
    }
    if (*c->s != '"') {
        c->err = "unterminated string";
        free(out);
        return NULL;
    }
    c->s++;
    out[len] = 0;
    return out;
}

static mj_node* parse_value(mj_cur* c);

static mj_node* parse_container(mj_cur* c, int isobj) {
    mj_node* n = node_new(isobj ? MJ_OBJECT : MJ_ARRAY);
    mj_node* last = NULL;
    char close = isobj ? '}' : ']';
    c->s++;
    if (++c->depth > 256) {
        c->err = "nesting too deep";
        return n;
    }
    skip_ws(c);
    if (*c->s == close) {
        c->s++;
        c->depth--;
        return n;
    }
    for (;;) {
        char* key = NULL;
        skip_ws(c);
        if (isobj) {
            key = parse_string_raw(c);
            if (!key) return n;
            skip_ws(c);
            if (*c->s != ':') {
                c->err = "expected ':'";
                free(key);
                return n;
            }
            c->s++;
        }
        mj_node* v = parse_value(c);
        if (!v) {
            free(key);
            return n;
        }
        v->key = key;
        if (last) last->next = v;
        else n->child = v;
        last = v;
        if (c->err) return n;
        skip_ws(c);
        if (*c->s == ',') {
            c->s++;
            continue;
        }
        if (*c->s == close) {
            c->s++;
            c->depth--;
            return n;
        }
        c->err = isobj ? "expected ',' or '}'" : "expected ',' or ']'";
        return n;
    }
}

static mj_node* parse_value(mj_cur* c) {
    skip_ws(c);
    const char* s = c->s;
    if (*s == '{' || *s == '[') return parse_container(c, *s == '{');
    if (*s == '"') {
        mj_node* n = node_new(MJ_STRING);
        n->str = parse_string_raw(c);
        return n;
    }
    if (!strncmp(s, "true", 4)) { c->s += 4; return node_new(MJ_TRUE); }
    if (!strncmp(s, "false", 5)) { c->s += 5; return node_new(MJ_FALSE); }
    if (!strncmp(s, "null", 4)) { c->s += 4; return node_new(MJ_NULL); }
    char* end = NULL;
    double v = strtod(s, &end);
    if (end == s) {
        c->err = "unexpected character";
        return NULL;
    }
    c->s = end;
    mj_node* n = node_new(MJ_NUMBER);
    n->num = v;
    return n;
}

mj_node* mj_parse(const char* text, char* err, int errlen) {
    mj_cur c = {text, NULL, 0};
    mj_node* root = parse_value(&c);
    if (!c.err) {
        skip_ws(&c);
        if (*c.s) c.err = "trailing characters";
    }
    if (c.err) {
        if (err && errlen > 0) snprintf(err, (size_t)errlen, "JSON error at offset %ld: %s", (long)(c.s - text), c.err);
        mj_free(root);
        return NULL;
    }
    return root;
}

void mj_free(mj_node* n) {
    while (n) {
        mj_node* nx = n->next;
        mj_free(n->child);
        free(n->str);
        free(n->key);
        free(n);
        n = nx;
    }
}

const mj_node* mj_get(const mj_node* obj, const char* key) {
    if (!obj || obj->type != MJ_OBJECT) return NULL;
    for (const mj_node* m = obj->child; m; m = m->next)
        if (m->key && !strcmp(m->key, key)) return m;
    return NULL;
}

const mj_node* mj_path(const mj_node* root, const char* path) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%s", path);
    const mj_node* n = root;
    for (char* tok = strtok(buf, "."); tok && n; tok = strtok(NULL, ".")) n = mj_get(n, tok);
    return n;
}

int mj_size(const mj_node* arr) {
    int k = 0;
    if (!arr || (arr->type != MJ_ARRAY && arr->type != MJ_OBJECT)) return 0;
    for (const mj_node* m = arr->child; m; m = m->next) k++;
    return k;
}

const mj_node* mj_at(const mj_node* arr, int i) {
    if (!arr) return NULL;
    const mj_node* m = arr->child;
    while (m && i-- > 0) m = m->next;
    return m;
}

double mj_num(const mj_node* n, double fallback) {
    if (!n) return fallback;
    if (n->type == MJ_NUMBER) return n->num;
    if (n->type == MJ_TRUE) return 1;
    if (n->type == MJ_FALSE) return 0;
    return fallback;
}

const char* mj_str(const mj_node* n, const char* fallback) {
    return (n && n->type == MJ_STRING) ? n->str : fallback;
}
