/*
 * mmc_json.h -- minimal JSON reader for the simulation input files
 * (objects, arrays, numbers, strings, true/false/null).
 */
#ifndef MMC_JSON_H
#define MMC_JSON_H

#ifdef __cplusplus
extern "C" {
#endif

enum { MJ_NULL, MJ_FALSE, MJ_TRUE, MJ_NUMBER, MJ_STRING, MJ_ARRAY, MJ_OBJECT };

typedef struct mj_node {
    int type;
    double num;
    char* str;              /* string value (MJ_STRING) */
    char* key;              /* member name when inside an object */
    struct mj_node* child;  /* first element / member */
    struct mj_node* next;   /* next sibling */
} mj_node;

/* Parses text; returns NULL and fills err (if given) on a syntax error. */
mj_node* mj_parse(const char* text, char* err, int errlen);
void mj_free(mj_node* n);
/* Member of an object by name, or NULL. */
const mj_node* mj_get(const mj_node* obj, const char* key);
/* Member at a dotted path such as "Optode.Source.Pos", or NULL. */
const mj_node* mj_path(const mj_node* root, const char* path);
int mj_size(const mj_node* arr);
const mj_node* mj_at(const mj_node* arr, int i);
double mj_num(const mj_node* n, double fallback);
const char* mj_str(const mj_node* n, const char* fallback);

#ifdef __cplusplus
}
#endif
#endif
