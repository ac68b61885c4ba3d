/* orc_util.h - small growable arrays for the oracle (test infrastructure). */
#ifndef ORC_UTIL_H
#define ORC_UTIL_H
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <math.h>

typedef struct { int64_t *a; int64_t len, cap; } ivec;

static inline void *orc_xmalloc(size_t sz) {
    void *p = malloc(sz ? sz : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}
static inline void *orc_xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz ? sz : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}
static inline void iv_push(ivec *v, int64_t x) {
    if (v->len == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 4;
        v->a = (int64_t *)realloc(v->a, (size_t)v->cap * sizeof(int64_t));
        if (!v->a) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    }
    v->a[v->len++] = x;
}
static inline int iv_contains(const ivec *v, int64_t x) {
    for (int64_t i = 0; i < v->len; i++) if (v->a[i] == x) return 1;
    return 0;
}
static inline void iv_free(ivec *v) { free(v->a); v->a = NULL; v->len = v->cap = 0; }

static int orc_cmp_i64(const void *x, const void *y) {
    int64_t a = *(const int64_t *)x, b = *(const int64_t *)y;
    return (a > b) - (a < b);
}
#endif
