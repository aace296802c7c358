/* Oracle MOS mapping: libsvm nu-SVR (svr.zig) and the TF-Lattice speech
 * model through a minimal TFLite flatbuffer interpreter (tflite.zig,
 * lattice.zig).  Test infrastructure only (see vo.h). */
#include "vo_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------
 * SVR (svr.zig:16-165)
 * ---------------------------------------------------------------------- */
struct vo_svr {
    double gamma, rho;
    size_t nf, nsv;
    double *sv, *coef;
};

static const char *line_end(const char *p, const char *end) {
    while (p < end && *p != '\n') p++;
    return p;
}

/* svr.zig:80-165 (parseModel) */
vo_svr *vo_svr_parse(const char *text, size_t len, int *err) {
    double gamma = 0.01, rho = 0.0;
    int saw_sv = 0;
    const char *p = text, *end = text + len;
    size_t cap = 64, nl = 0, max_feat = 0;
    const char **ls = (const char **)malloc(sizeof(char *) * cap);
    size_t *ll = (size_t *)malloc(sizeof(size_t) * cap);
    *err = VO_OK;
    while (p <= end) {
        const char *e = line_end(p, end);
        size_t n = (size_t)(e - p);
        if (n > 0 && p[n - 1] == '\r') n--;
        if (n > 0) {
            if (!saw_sv) {
                if (n == 2 && p[0] == 'S' && p[1] == 'V') {
                    saw_sv = 1;
                } else {
                    const char *sp = memchr(p, ' ', n);
                    if (sp) {
                        size_t kl = (size_t)(sp - p);
                        char val[128];
                        const char *v = sp + 1, *ve = p + n;
                        while (v < ve && (*v == ' ' || *v == '\t')) v++;
                        while (ve > v && (ve[-1] == ' ' || ve[-1] == '\t')) ve--;
                        size_t vl = (size_t)(ve - v) < 127 ? (size_t)(ve - v) : 127;
                        memcpy(val, v, vl);
                        val[vl] = 0;
                        if (kl == 8 && !memcmp(p, "svm_type", 8)) {
                            if (strcmp(val, "nu_svr") && strcmp(val, "epsilon_svr")) { *err = VO_E_MODEL; goto fail; }
                        } else if (kl == 11 && !memcmp(p, "kernel_type", 11)) {
                            if (strcmp(val, "rbf")) { *err = VO_E_MODEL; goto fail; }
                        } else if (kl == 5 && !memcmp(p, "gamma", 5)) {
                            char *ep;
                            gamma = strtod(val, &ep);
                            if (ep == val || *ep) { *err = VO_E_MODEL; goto fail; }
                        } else if (kl == 3 && !memcmp(p, "rho", 3)) {
                            char *ep;
                            rho = strtod(val, &ep);
                            if (ep == val || *ep) { *err = VO_E_MODEL; goto fail; }
                        }
                    }
                }
            } else {
                if (nl == cap) {
                    cap *= 2;
                    ls = (const char **)realloc(ls, sizeof(char *) * cap);
                    ll = (size_t *)realloc(ll, sizeof(size_t) * cap);
                }
                ls[nl] = p;
                ll[nl] = n;
                nl++;
                /* max feature index over "idx:val" tokens after the first */
                const char *q = p, *qe = p + n;
                int first = 1;
                while (q < qe) {
                    while (q < qe && (*q == ' ' || *q == '\t')) q++;
                    const char *ts = q;
                    while (q < qe && *q != ' ' && *q != '\t') q++;
                    if (q == ts) break;
                    if (first) { first = 0; continue; }
                    const char *colon = memchr(ts, ':', (size_t)(q - ts));
                    if (!colon || colon == ts) continue;
                    size_t idx = 0;
                    int ok = 1;
                    for (const char *d = ts; d < colon; d++) {
                        if (*d < '0' || *d > '9') { ok = 0; break; }
                        idx = idx * 10 + (size_t)(*d - '0');
                    }
                    if (ok && idx > max_feat) max_feat = idx;
                }
            }
        }
        if (e >= end) break;
        p = e + 1;
    }
    if (nl == 0 || max_feat == 0) { *err = VO_E_MODEL; goto fail; }
    vo_svr *m = (vo_svr *)calloc(1, sizeof(vo_svr));
    m->gamma = gamma; m->rho = rho; m->nf = max_feat; m->nsv = nl;
    m->sv = (double *)calloc(nl * max_feat, sizeof(double));
    m->coef = (double *)malloc(sizeof(double) * nl);
    for (size_t s = 0; s < nl; s++) {
        const char *q = ls[s], *qe = ls[s] + ll[s];
        int first = 1;
        while (q < qe) {
            while (q < qe && (*q == ' ' || *q == '\t')) q++;
            const char *ts = q;
            while (q < qe && *q != ' ' && *q != '\t') q++;
            if (q == ts) break;
            char tok[128];
            size_t tl = (size_t)(q - ts) < 127 ? (size_t)(q - ts) : 127;
            memcpy(tok, ts, tl);
            tok[tl] = 0;
            if (first) {
                char *ep;
                m->coef[s] = strtod(tok, &ep);
                if (ep == tok || *ep) { *err = VO_E_MODEL; vo_svr_free(m); goto fail; }
                first = 0;
                continue;
            }
            char *colon = strchr(tok, ':');
            if (!colon) continue;
            *colon = 0;
            char *ep;
            unsigned long long idx = strtoull(tok, &ep, 10);
            if (ep == tok || *ep) continue;
            double v = strtod(colon + 1, &ep);
            if (ep == colon + 1 || *ep) continue;
            if (idx == 0 || idx > max_feat) { *err = VO_E_MODEL; vo_svr_free(m); goto fail; }
            m->sv[s * max_feat + (idx - 1)] = v;
        }
    }
    free(ls); free(ll);
    return m;
fail:
    free(ls); free(ll);
    return NULL;
}

void vo_svr_free(vo_svr *m) {
    if (!m) return;
    free(m->sv); free(m->coef); free(m);
}

/* svr.zig:37-61 (predict, f64 kernel) */
double vo_svr_eval(const vo_svr *m, const double *obs, size_t n) {
    (void)n;
    double sum = 0.0;
    for (size_t i = 0; i < m->nsv; i++) {
        const double *sv = m->sv + i * m->nf;
        double dist = 0.0;
        for (size_t k = 0; k < m->nf; k++) {
            double d = sv[k] - obs[k];
            dist += d * d;
        }
        sum += m->coef[i] * exp(-m->gamma * dist);
    }
    return sum - m->rho;
}

int vo_svr_predict(const char *model, size_t len, const double *obs, size_t n, double *y) {
    int err;
    vo_svr *m = vo_svr_parse(model, len, &err);
    if (!m) return err;
    if (n != m->nf) { vo_svr_free(m); return VO_E_INVALID_ARGUMENT; }
    *y = vo_svr_eval(m, obs, n);
    vo_svr_free(m);
    return VO_OK;
}

/* ------------------------------------------------------------------------
 * TFLite flatbuffer access (tflite.zig:15-82)
 * ---------------------------------------------------------------------- */
typedef struct { const uint8_t *buf; size_t len; } fb;
typedef struct { size_t pos; } tbl;
typedef struct { size_t start, len; } vec;

static int rd(const fb *b, size_t pos, size_t sz, uint64_t *out) {
    if (pos > b->len || b->len - pos < sz) return -1;
    uint64_t v = 0;
    for (size_t i = 0; i < sz; i++) v |= (uint64_t)b->buf[pos + i] << (8 * i);
    *out = v;
    return 0;
}

/* absolute position of field id, 0 when absent, (size_t)-1 on error */
static size_t fb_field(const fb *b, tbl t, unsigned id) {
    uint64_t soff_u, vts, off;
    if (rd(b, t.pos, 4, &soff_u)) return (size_t)-1;
    int64_t vt = (int64_t)t.pos - (int64_t)(int32_t)(uint32_t)soff_u;
    if (vt < 0) return (size_t)-1;
    if (rd(b, (size_t)vt, 2, &vts)) return (size_t)-1;
    size_t slot = 4 + 2 * (size_t)id;
    if (slot + 2 > vts) return 0;
    if (rd(b, (size_t)vt + slot, 2, &off)) return (size_t)-1;
    return off == 0 ? 0 : t.pos + (size_t)off;
}

static int fb_scalar(const fb *b, tbl t, unsigned id, size_t sz, uint64_t dflt, uint64_t *out) {
    size_t p = fb_field(b, t, id);
    if (p == (size_t)-1) return -1;
    if (p == 0) { *out = dflt; return 0; }
    return rd(b, p, sz, out);
}

static int fb_deref(const fb *b, size_t p, size_t *out) {
    uint64_t v;
    if (rd(b, p, 4, &v)) return -1;
    *out = p + (size_t)v;
    return 0;
}

/* 1 present, 0 absent, -1 error */
static int fb_table(const fb *b, tbl t, unsigned id, tbl *out) {
    size_t p = fb_field(b, t, id);
    if (p == (size_t)-1) return -1;
    if (p == 0) return 0;
    if (fb_deref(b, p, &out->pos)) return -1;
    return 1;
}

static int fb_vector(const fb *b, tbl t, unsigned id, vec *out) {
    size_t p = fb_field(b, t, id);
    if (p == (size_t)-1) return -1;
    if (p == 0) return 0;
    size_t v;
    uint64_t n;
    if (fb_deref(b, p, &v) || rd(b, v, 4, &n)) return -1;
    out->start = v + 4;
    out->len = (size_t)n;
    return 1;
}

static int vec_table(const fb *b, vec v, size_t i, tbl *out) { return fb_deref(b, v.start + 4 * i, &out->pos); }
static int vec_i32(const fb *b, vec v, size_t i, int32_t *out) {
    uint64_t x;
    if (rd(b, v.start + 4 * i, 4, &x)) return -1;
    *out = (int32_t)(uint32_t)x;
    return 0;
}
static int fb_string(const fb *b, tbl t, unsigned id, const char **s, size_t *n) {
    vec v;
    int r = fb_vector(b, t, id, &v);
    if (r <= 0) { *s = ""; *n = 0; return r; }
    if (v.start > b->len || b->len - v.start < v.len) return -1;
    *s = (const char *)b->buf + v.start;
    *n = v.len;
    return 1;
}

/* ------------------------------------------------------------------------
 * Tensors and ops (tflite.zig:88-212)
 * ---------------------------------------------------------------------- */
enum { DT_NONE = 0, DT_F32, DT_I32, DT_BOOL };
#define MAXR 6
typedef struct { int rank; size_t d[MAXR]; } shp;
typedef struct { int set; shp s; int dt; void *data; } val;

static size_t shp_count(shp s) { size_t n = 1; for (int k = 0; k < s.rank; k++) n *= s.d[k]; return n; }

enum { K_ADD, K_SUB, K_MUL, K_MIN, K_MAX, K_EQ, K_RELU, K_RELU6, K_ABS, K_CAST, K_CONCAT, K_FC, K_BMM, K_RESHAPE,
       K_EXPAND, K_SPLIT, K_SHAPE, K_FILL, K_UNSUPPORTED };

static int kind_of(int32_t code) { /* tflite.zig:157-179 */
    switch (code) {
        case 0: return K_ADD; case 2: return K_CONCAT; case 9: return K_FC; case 18: return K_MUL;
        case 19: return K_RELU; case 21: return K_RELU6; case 22: return K_RESHAPE; case 41: return K_SUB;
        case 49: return K_SPLIT; case 53: return K_CAST; case 55: return K_MAX; case 57: return K_MIN;
        case 70: return K_EXPAND; case 71: return K_EQ; case 77: return K_SHAPE; case 94: return K_FILL;
        case 101: return K_ABS; case 126: return K_BMM; default: return K_UNSUPPORTED;
    }
}

enum { ACT_NONE, ACT_RELU, ACT_N1_1, ACT_RELU6 };
static float activate(int act, float x) { /* tflite.zig:193-200 */
    switch (act) {
        case ACT_RELU: return fmaxf(x, 0.0f);
        case ACT_N1_1: return fmaxf(-1.0f, fminf(x, 1.0f));
        case ACT_RELU6: return fmaxf(0.0f, fminf(x, 6.0f));
        default: return x;
    }
}

typedef struct {
    int kind, act;
    int32_t axis;
    int adj_x, adj_y, keep;
    int32_t *in, *out;
    size_t nin, nout;
    int32_t *new_shape;
    size_t n_new_shape;
    int has_new_shape;
} op_t;

typedef struct { const char *name; size_t nlen; uint32_t tensor; } named;

struct vo_lattice {
    size_t ntens, nops, nin, nout, nb;
    int *dtypes;
    val *consts;
    op_t *ops;
    named *inputs, *outputs;
    uint32_t feat[4 * VO_MAX_BANDS], tau, output;
    void **arena;
    size_t narena, carena;
    uint8_t *bytes;
};

static void *lat_alloc(struct vo_lattice *m, size_t n) {
    void *p = calloc(1, n ? n : 1);
    if (m->narena == m->carena) {
        m->carena = m->carena ? 2 * m->carena : 256;
        m->arena = (void **)realloc(m->arena, sizeof(void *) * m->carena);
    }
    m->arena[m->narena++] = p;
    return p;
}

/* scratch allocations for one run */
typedef struct { void **p; size_t n, cap; } scratch;
static void *sc_alloc(scratch *s, size_t n) {
    if (s->n == s->cap) { s->cap = s->cap ? s->cap * 2 : 1024; s->p = (void **)realloc(s->p, sizeof(void *) * s->cap); }
    void *p = calloc(1, n ? n : 1);
    s->p[s->n++] = p;
    return p;
}
static void sc_free(scratch *s) { for (size_t i = 0; i < s->n; i++) free(s->p[i]); free(s->p); }

static int dtype_of(int t) { return t == 0 ? DT_F32 : t == 2 ? DT_I32 : t == 6 ? DT_BOOL : DT_NONE; }

static int tensor_list(struct vo_lattice *m, const fb *b, tbl t, unsigned id, int32_t **out, size_t *n) {
    vec v;
    int r = fb_vector(b, t, id, &v);
    if (r < 0) return -1;
    if (r == 0) { *out = NULL; *n = 0; return 0; }
    *out = (int32_t *)lat_alloc(m, sizeof(int32_t) * v.len);
    *n = v.len;
    for (size_t i = 0; i < v.len; i++) {
        if (vec_i32(b, v, i, &(*out)[i])) return -1;
        if ((*out)[i] >= (int64_t)m->ntens) return -1;
    }
    return 0;
}

static int tensor_maps(struct vo_lattice *m, const fb *b, tbl t, unsigned id, named **out, size_t *n) {
    vec v;
    int r = fb_vector(b, t, id, &v);
    if (r < 0) return -1;
    if (r == 0) { *out = NULL; *n = 0; return 0; }
    *out = (named *)lat_alloc(m, sizeof(named) * v.len);
    *n = v.len;
    for (size_t i = 0; i < v.len; i++) {
        tbl e;
        uint64_t ti;
        if (vec_table(b, v, i, &e)) return -1;
        if (fb_string(b, e, 0, &(*out)[i].name, &(*out)[i].nlen) < 0) return -1;
        if (fb_scalar(b, e, 1, 4, 0, &ti)) return -1;
        (*out)[i].tensor = (uint32_t)ti;
        if (ti >= m->ntens) return -1;
    }
    return 0;
}

static int find_named(const named *l, size_t n, const char *name, uint32_t *t) {
    size_t nl = strlen(name);
    for (size_t i = 0; i < n; i++)
        if (l[i].nlen == nl && !memcmp(l[i].name, name, nl)) { *t = l[i].tensor; return 1; }
    return 0;
}

/* tflite.zig:232-360 (Model.init) + lattice.zig:23-41 (Mapper.init) */
vo_lattice *vo_lattice_load(const uint8_t *bytes, size_t len, size_t nb, int *err) {
    struct vo_lattice *m = (struct vo_lattice *)calloc(1, sizeof(struct vo_lattice));
    fb b = {bytes, len};
    *err = VO_E_MODEL;
    m->nb = nb;
    uint64_t root_pos;
    if (rd(&b, 0, 4, &root_pos)) goto fail;
    tbl root = {(size_t)root_pos};
    vec codes_v, subgraphs, buffers, tensors_v, ops_v;
    if (fb_vector(&b, root, 1, &codes_v) != 1) goto fail;
    if (fb_vector(&b, root, 2, &subgraphs) != 1) goto fail;
    if (fb_vector(&b, root, 4, &buffers) != 1) goto fail;
    if (subgraphs.len < 1) goto fail;
    tbl g;
    if (vec_table(&b, subgraphs, 0, &g)) goto fail;
    int *codes = (int *)lat_alloc(m, sizeof(int) * codes_v.len);
    for (size_t i = 0; i < codes_v.len; i++) {
        tbl t;
        uint64_t dep, bi;
        if (vec_table(&b, codes_v, i, &t)) goto fail;
        size_t cf = fb_field(&b, t, 1);
        if (cf == (size_t)-1 || cf != 0) goto fail; /* custom op */
        if (fb_scalar(&b, t, 0, 1, 0, &dep) || fb_scalar(&b, t, 3, 4, 0, &bi)) goto fail;
        int32_t d = (int32_t)(int8_t)(uint8_t)dep, bc = (int32_t)(uint32_t)bi;
        codes[i] = kind_of(d > bc ? d : bc);
    }
    if (fb_vector(&b, g, 0, &tensors_v) != 1) goto fail;
    m->ntens = tensors_v.len;
    const char **names = (const char **)lat_alloc(m, sizeof(char *) * m->ntens);
    size_t *name_lens = (size_t *)lat_alloc(m, sizeof(size_t) * m->ntens);
    m->dtypes = (int *)lat_alloc(m, sizeof(int) * m->ntens);
    m->consts = (val *)lat_alloc(m, sizeof(val) * m->ntens);
    for (size_t i = 0; i < m->ntens; i++) {
        tbl t;
        uint64_t ty, bidx;
        if (vec_table(&b, tensors_v, i, &t)) goto fail;
        if (fb_string(&b, t, 3, &names[i], &name_lens[i]) < 0) goto fail;
        if (fb_scalar(&b, t, 1, 1, 0, &ty)) goto fail;
        m->dtypes[i] = dtype_of((int)(int8_t)(uint8_t)ty);
        if (fb_scalar(&b, t, 2, 4, 0, &bidx)) goto fail;
        if (bidx == 0 || bidx >= buffers.len) continue;
        tbl bt;
        if (vec_table(&b, buffers, (size_t)bidx, &bt)) goto fail;
        vec dv;
        const uint8_t *data = NULL;
        size_t dlen = 0;
        int r = fb_vector(&b, bt, 0, &dv);
        if (r < 0) goto fail;
        if (r == 1) {
            if (dv.start > len || len - dv.start < dv.len) goto fail;
            data = bytes + dv.start;
            dlen = dv.len;
        }
        if (dlen == 0) {
            uint64_t off, size;
            if (fb_scalar(&b, bt, 1, 8, 0, &off) || fb_scalar(&b, bt, 2, 8, 0, &size)) goto fail;
            if (off > 1 && size > 0) {
                if (off > len || len - off < size) goto fail;
                data = bytes + off;
                dlen = (size_t)size;
            }
        }
        if (dlen == 0) continue;
        vec sv;
        int rs = fb_vector(&b, t, 0, &sv);
        if (rs < 0) goto fail;
        shp s = {0, {0}};
        if (rs == 1) {
            if (sv.len > MAXR) { *err = VO_E_MODEL; goto fail; }
            s.rank = (int)sv.len;
            for (size_t d = 0; d < sv.len; d++) {
                int32_t x;
                if (vec_i32(&b, sv, d, &x) || x < 0) goto fail;
                s.d[d] = (size_t)x;
            }
        }
        size_t n = shp_count(s);
        int dt = m->dtypes[i];
        if (dt == DT_NONE) goto fail;
        val *c = &m->consts[i];
        c->set = 1; c->s = s; c->dt = dt;
        if (dt == DT_BOOL) {
            if (dlen != n) goto fail;
            uint8_t *o = (uint8_t *)lat_alloc(m, n);
            for (size_t k = 0; k < n; k++) o[k] = data[k] != 0;
            c->data = o;
        } else {
            if (dlen != n * 4) goto fail;
            uint32_t *o = (uint32_t *)lat_alloc(m, n * 4);
            for (size_t k = 0; k < n; k++)
                o[k] = (uint32_t)data[4 * k] | (uint32_t)data[4 * k + 1] << 8 | (uint32_t)data[4 * k + 2] << 16 |
                       (uint32_t)data[4 * k + 3] << 24;
            c->data = o;
        }
    }
    if (fb_vector(&b, g, 3, &ops_v) != 1) goto fail;
    m->nops = ops_v.len;
    m->ops = (op_t *)lat_alloc(m, sizeof(op_t) * m->nops);
    for (size_t i = 0; i < m->nops; i++) {
        tbl t;
        uint64_t ci;
        if (vec_table(&b, ops_v, i, &t)) goto fail;
        if (fb_scalar(&b, t, 0, 4, 0, &ci)) goto fail;
        if (ci >= codes_v.len) goto fail;
        op_t *op = &m->ops[i];
        op->kind = codes[ci];
        if (op->kind == K_UNSUPPORTED) goto fail;
        if (tensor_list(m, &b, t, 1, &op->in, &op->nin) || tensor_list(m, &b, t, 2, &op->out, &op->nout)) goto fail;
        for (size_t k = 0; k < op->nout; k++) if (op->out[k] < 0) goto fail;
        tbl o;
        int ro = fb_table(&b, t, 4, &o);
        if (ro < 0) goto fail;
        uint64_t x;
        if (ro == 1) {
            switch (op->kind) {
                case K_ADD: case K_SUB: case K_MUL:
                    if (fb_scalar(&b, o, 0, 1, 0, &x) || (int8_t)x < 0 || (int8_t)x > 3) goto fail;
                    op->act = (int)(int8_t)x;
                    break;
                case K_CONCAT:
                    if (fb_scalar(&b, o, 0, 4, 0, &x)) goto fail;
                    op->axis = (int32_t)(uint32_t)x;
                    if (fb_scalar(&b, o, 1, 1, 0, &x) || (int8_t)x < 0 || (int8_t)x > 3) goto fail;
                    op->act = (int)(int8_t)x;
                    break;
                case K_FC:
                    if (fb_scalar(&b, o, 0, 1, 0, &x) || (int8_t)x < 0 || (int8_t)x > 3) goto fail;
                    op->act = (int)(int8_t)x;
                    if (fb_scalar(&b, o, 1, 1, 0, &x) || x != 0) goto fail;
                    if (fb_scalar(&b, o, 2, 1, 0, &x)) goto fail;
                    op->keep = x != 0;
                    break;
                case K_BMM:
                    if (fb_scalar(&b, o, 0, 1, 0, &x)) goto fail;
                    op->adj_x = x != 0;
                    if (fb_scalar(&b, o, 1, 1, 0, &x)) goto fail;
                    op->adj_y = x != 0;
                    break;
                case K_RESHAPE: {
                    vec nv;
                    int rv = fb_vector(&b, o, 0, &nv);
                    if (rv < 0) goto fail;
                    if (rv == 1) {
                        op->has_new_shape = 1;
                        op->n_new_shape = nv.len;
                        op->new_shape = (int32_t *)lat_alloc(m, sizeof(int32_t) * nv.len);
                        for (size_t d = 0; d < nv.len; d++) if (vec_i32(&b, nv, d, &op->new_shape[d])) goto fail;
                    }
                    break;
                }
                default: break;
            }
        }
    }
    vec sigs;
    int rsig = fb_vector(&b, root, 7, &sigs);
    if (rsig < 0) goto fail;
    if (rsig == 1) {
        for (size_t i = 0; i < sigs.len; i++) {
            tbl s;
            const char *key;
            size_t kl;
            uint64_t sgi;
            if (vec_table(&b, sigs, i, &s) || fb_string(&b, s, 2, &key, &kl) < 0) goto fail;
            if (!(kl == 7 && !memcmp(key, "predict", 7)) && sigs.len > 1) continue;
            if (fb_scalar(&b, s, 4, 4, 0, &sgi) || sgi != 0) goto fail;
            if (tensor_maps(m, &b, s, 0, &m->inputs, &m->nin) || tensor_maps(m, &b, s, 1, &m->outputs, &m->nout))
                goto fail;
            break;
        }
    }
    if (m->nin == 0) {
        vec iv, ov;
        if (fb_vector(&b, g, 1, &iv) < 0 || fb_vector(&b, g, 2, &ov) < 0) goto fail;
        m->inputs = (named *)lat_alloc(m, sizeof(named) * (iv.len + 1));
        m->outputs = (named *)lat_alloc(m, sizeof(named) * (ov.len + 1));
        for (size_t i = 0; i < iv.len; i++) {
            int32_t ti;
            if (vec_i32(&b, iv, i, &ti) || ti < 0 || (size_t)ti >= m->ntens) goto fail;
            m->inputs[i] = (named){names[ti], name_lens[ti], (uint32_t)ti};
        }
        for (size_t i = 0; i < ov.len; i++) {
            int32_t ti;
            if (vec_i32(&b, ov, i, &ti) || ti < 0 || (size_t)ti >= m->ntens) goto fail;
            m->outputs[i] = (named){names[ti], name_lens[ti], (uint32_t)ti};
        }
        m->nin = iv.len;
        m->nout = ov.len;
    }
    static const char *FEATURES[4] = {"fvnsim", "fvnsim10_", "fstdnsim", "fvdegenergy"};
    char nm[64];
    for (int fi = 0; fi < 4; fi++)
        for (size_t bnd = 0; bnd < nb; bnd++) {
            snprintf(nm, sizeof nm, "%s%zu", FEATURES[fi], bnd);
            if (!find_named(m->inputs, m->nin, nm, &m->feat[fi * nb + bnd])) goto fail;
        }
    if (!find_named(m->inputs, m->nin, "tau", &m->tau)) goto fail;
    if (!find_named(m->outputs, m->nout, "predictions", &m->output)) goto fail;
    *err = VO_OK;
    return m;
fail:
    vo_lattice_free(m);
    return NULL;
}

void vo_lattice_free(vo_lattice *m) {
    if (!m) return;
    for (size_t i = 0; i < m->narena; i++) free(m->arena[i]);
    free(m->arena);
    free(m);
}

/* ---- kernels (tflite.zig:464-793) ---- */
static int broadcast(shp x, shp y, shp *out) {
    out->rank = x.rank > y.rank ? x.rank : y.rank;
    for (int k = 0; k < out->rank; k++) {
        size_t dx = k + x.rank >= out->rank ? x.d[k + x.rank - out->rank] : 1;
        size_t dy = k + y.rank >= out->rank ? y.d[k + y.rank - out->rank] : 1;
        if (dx != dy && dx != 1 && dy != 1) return -1;
        out->d[k] = dx == 1 ? dy : dx;
    }
    return 0;
}

static void strides_of(shp s, shp out, size_t *st) {
    for (int k = 0; k < MAXR; k++) st[k] = 0;
    size_t acc = 1;
    for (int k = s.rank - 1; k >= 0; k--) {
        int ok = k + out.rank - s.rank;
        st[ok] = s.d[k] == 1 ? 0 : acc;
        acc *= s.d[k];
    }
}

/* Visits broadcast shape `out` in row-major order, yielding x / y offsets. */
typedef struct { shp out; size_t sx[MAXR], sy[MAXR], idx[MAXR], ix, iy; } bcast_it;
static void bc_init(bcast_it *it, shp out, shp x, shp y) {
    it->out = out;
    strides_of(x, out, it->sx);
    strides_of(y, out, it->sy);
    memset(it->idx, 0, sizeof it->idx);
    it->ix = it->iy = 0;
}
static void bc_next(bcast_it *it) {
    for (int k = it->out.rank - 1; k >= 0; k--) {
        it->idx[k]++;
        it->ix += it->sx[k];
        it->iy += it->sy[k];
        if (it->idx[k] < it->out.d[k]) return;
        it->ix -= it->sx[k] * it->out.d[k];
        it->iy -= it->sy[k] * it->out.d[k];
        it->idx[k] = 0;
    }
}

static int get_in(val *vals, const op_t *op, size_t i, val *out) {
    if (i >= op->nin || op->in[i] < 0) return -1;
    *out = vals[op->in[i]];
    return out->set ? 0 : -1;
}

static int norm_axis(int32_t axis, int rank, int *out) {
    int ax = axis < 0 ? axis + rank : axis;
    if (ax < 0 || ax >= rank) return -1;
    *out = ax;
    return 0;
}

static float dot_fma(const float *x, const float *w, size_t n) { /* tflite.zig:667-671 */
    float acc = 0.0f;
    for (size_t i = 0; i < n; i++) acc = fmaf(x[i], w[i], acc);
    return acc;
}

static int exec_op(const struct vo_lattice *m, scratch *sc, const op_t *op, val *vals) {
    if (op->nout == 0) return -1;
    val x, y, r = {1, {0, {0}}, 0, NULL};
    size_t out0 = (size_t)op->out[0];
    switch (op->kind) {
        case K_ADD: case K_SUB: case K_MUL: case K_MIN: case K_MAX: {
            if (get_in(vals, op, 0, &x) || get_in(vals, op, 1, &y)) return -1;
            if (x.dt != DT_F32 || y.dt != DT_F32) return -1;
            if (broadcast(x.s, y.s, &r.s)) return -1;
            size_t n = shp_count(r.s);
            float *o = (float *)sc_alloc(sc, 4 * n);
            const float *xs = (const float *)x.data, *ys = (const float *)y.data;
            bcast_it it;
            bc_init(&it, r.s, x.s, y.s);
            for (size_t i = 0; i < n; i++) {
                float p = xs[it.ix], q = ys[it.iy], v;
                switch (op->kind) {
                    case K_ADD: v = p + q; break;
                    case K_SUB: v = p - q; break;
                    case K_MUL: v = p * q; break;
                    case K_MIN: v = fminf(p, q); break;
                    default: v = fmaxf(p, q); break;
                }
                o[i] = activate(op->act, v);
                bc_next(&it);
            }
            r.dt = DT_F32; r.data = o;
            break;
        }
        case K_EQ: {
            if (get_in(vals, op, 0, &x) || get_in(vals, op, 1, &y)) return -1;
            if (broadcast(x.s, y.s, &r.s)) return -1;
            if (x.dt == DT_BOOL || x.dt != y.dt) return -1;
            size_t n = shp_count(r.s);
            uint8_t *o = (uint8_t *)sc_alloc(sc, n);
            bcast_it it;
            bc_init(&it, r.s, x.s, y.s);
            for (size_t i = 0; i < n; i++) {
                if (x.dt == DT_F32) o[i] = ((const float *)x.data)[it.ix] == ((const float *)y.data)[it.iy];
                else o[i] = ((const int32_t *)x.data)[it.ix] == ((const int32_t *)y.data)[it.iy];
                bc_next(&it);
            }
            r.dt = DT_BOOL; r.data = o;
            break;
        }
        case K_RELU: case K_RELU6: case K_ABS: {
            if (get_in(vals, op, 0, &x) || x.dt != DT_F32) return -1;
            size_t n = shp_count(x.s);
            float *o = (float *)sc_alloc(sc, 4 * n);
            const float *xs = (const float *)x.data;
            for (size_t i = 0; i < n; i++)
                o[i] = op->kind == K_RELU ? fmaxf(xs[i], 0.0f) : op->kind == K_RELU6 ? fmaxf(0.0f, fminf(xs[i], 6.0f)) : fabsf(xs[i]);
            r.s = x.s; r.dt = DT_F32; r.data = o;
            break;
        }
        case K_CAST: {
            if (get_in(vals, op, 0, &x)) return -1;
            int to = m->dtypes[out0];
            if (to == DT_NONE) return -1;
            size_t n = shp_count(x.s);
            void *o = sc_alloc(sc, (to == DT_BOOL ? 1 : 4) * n);
            for (size_t i = 0; i < n; i++) {
                double v; /* exact for every f32 / i32 / bool source */
                if (x.dt == DT_F32) v = ((const float *)x.data)[i];
                else if (x.dt == DT_I32) v = ((const int32_t *)x.data)[i];
                else v = ((const uint8_t *)x.data)[i];
                if (to == DT_F32) {
                    ((float *)o)[i] = x.dt == DT_F32 ? ((const float *)x.data)[i] : (float)v;
                } else if (to == DT_I32) {
                    if (x.dt == DT_I32) ((int32_t *)o)[i] = ((const int32_t *)x.data)[i];
                    else {
                        double t = trunc(v);
                        ((int32_t *)o)[i] = isnan(t) ? 0 : t >= 2147483647.0 ? 2147483647 : t <= -2147483648.0 ? (int32_t)-2147483647 - 1 : (int32_t)t;
                    }
                } else {
                    ((uint8_t *)o)[i] = v != 0;
                }
            }
            r.s = x.s; r.dt = to; r.data = o;
            break;
        }
        case K_CONCAT: {
            if (op->nin == 0) return -1;
            val first = vals[op->in[0]];
            if (!first.set) return -1;
            int ax;
            if (norm_axis(op->axis, first.s.rank, &ax)) return -1;
            r.s = first.s;
            r.s.d[ax] = 0;
            for (size_t i = 0; i < op->nin; i++) {
                val v = vals[op->in[i]];
                if (!v.set || v.s.rank != r.s.rank) return -1;
                for (int k = 0; k < r.s.rank; k++) if (k != ax && v.s.d[k] != first.s.d[k]) return -1;
                r.s.d[ax] += v.s.d[ax];
            }
            size_t outer = 1, inner = 1;
            for (int k = 0; k < ax; k++) outer *= r.s.d[k];
            for (int k = ax + 1; k < r.s.rank; k++) inner *= r.s.d[k];
            float *o = (float *)sc_alloc(sc, 4 * shp_count(r.s));
            size_t pos = 0;
            for (size_t oi = 0; oi < outer; oi++)
                for (size_t i = 0; i < op->nin; i++) {
                    val v = vals[op->in[i]];
                    if (v.dt != DT_F32) return -1;
                    size_t chunk = v.s.d[ax] * inner;
                    const float *xs = (const float *)v.data + oi * chunk;
                    for (size_t k = 0; k < chunk; k++) o[pos + k] = activate(op->act, xs[k]);
                    pos += chunk;
                }
            r.dt = DT_F32; r.data = o;
            break;
        }
        case K_FC: {
            val w, bias = {0};
            int has_bias = 0;
            if (get_in(vals, op, 0, &x) || get_in(vals, op, 1, &w)) return -1;
            if (op->nin > 2 && op->in[2] >= 0) { if (get_in(vals, op, 2, &bias)) return -1; has_bias = 1; }
            if (x.dt != DT_F32 || w.dt != DT_F32 || w.s.rank != 2) return -1;
            size_t n_out = w.s.d[0], n_in = w.s.d[1], xl = shp_count(x.s);
            if (n_in == 0 || xl % n_in) return -1;
            size_t batch = xl / n_in;
            if (has_bias && (bias.dt != DT_F32 || shp_count(bias.s) != n_out)) return -1;
            float *o = (float *)sc_alloc(sc, 4 * batch * n_out);
            for (size_t rr = 0; rr < batch; rr++)
                for (size_t oo = 0; oo < n_out; oo++) {
                    float v = dot_fma((const float *)x.data + rr * n_in, (const float *)w.data + oo * n_in, n_in);
                    if (has_bias) v += ((const float *)bias.data)[oo];
                    o[rr * n_out + oo] = activate(op->act, v);
                }
            if (op->keep) { r.s = x.s; r.s.d[r.s.rank - 1] = n_out; }
            else { r.s.rank = 2; r.s.d[0] = batch; r.s.d[1] = n_out; }
            r.dt = DT_F32; r.data = o;
            break;
        }
        case K_BMM: {
            if (get_in(vals, op, 0, &x) || get_in(vals, op, 1, &y)) return -1;
            if (x.dt != DT_F32 || y.dt != DT_F32 || x.s.rank < 2 || y.s.rank < 2) return -1;
            int xr = x.s.rank, yr = y.s.rank;
            size_t mm = op->adj_x ? x.s.d[xr - 1] : x.s.d[xr - 2];
            size_t kk = op->adj_x ? x.s.d[xr - 2] : x.s.d[xr - 1];
            size_t k2 = op->adj_y ? y.s.d[yr - 1] : y.s.d[yr - 2];
            size_t nn = op->adj_y ? y.s.d[yr - 2] : y.s.d[yr - 1];
            if (kk != k2) return -1;
            shp xb = {xr - 2, {0}}, yb = {yr - 2, {0}}, ob;
            for (int k = 0; k < xr - 2; k++) xb.d[k] = x.s.d[k];
            for (int k = 0; k < yr - 2; k++) yb.d[k] = y.s.d[k];
            if (broadcast(xb, yb, &ob)) return -1;
            size_t nb_ = shp_count(ob);
            float *o = (float *)sc_alloc(sc, 4 * nb_ * mm * nn);
            bcast_it it;
            bc_init(&it, ob, xb, yb);
            const float *xs = (const float *)x.data, *ys = (const float *)y.data;
            for (size_t bi = 0; bi < nb_; bi++) {
                const float *xm = xs + it.ix * mm * kk, *ym = ys + it.iy * kk * nn;
                float *om = o + bi * mm * nn;
                for (size_t i = 0; i < mm; i++)
                    for (size_t j = 0; j < nn; j++) {
                        float acc = 0.0f;
                        for (size_t l = 0; l < kk; l++) {
                            float p = op->adj_x ? xm[l * mm + i] : xm[i * kk + l];
                            float q = op->adj_y ? ym[j * kk + l] : ym[l * nn + j];
                            acc = fmaf(p, q, acc);
                        }
                        om[i * nn + j] = acc;
                    }
                bc_next(&it);
            }
            if (ob.rank + 2 > MAXR) return -1;
            r.s = ob;
            r.s.d[r.s.rank] = mm;
            r.s.d[r.s.rank + 1] = nn;
            r.s.rank += 2;
            r.dt = DT_F32; r.data = o;
            break;
        }
        case K_RESHAPE: {
            if (get_in(vals, op, 0, &x)) return -1;
            const int32_t *spec;
            size_t ns;
            val sv;
            if (op->nin > 1 && op->in[1] >= 0) {
                if (get_in(vals, op, 1, &sv) || sv.dt != DT_I32) return -1;
                spec = (const int32_t *)sv.data;
                ns = shp_count(sv.s);
            } else if (op->has_new_shape) {
                spec = op->new_shape;
                ns = op->n_new_shape;
            } else return -1;
            if (ns > MAXR) return -1;
            size_t total = shp_count(x.s), known = 1;
            int infer = -1;
            r.s.rank = (int)ns;
            for (size_t k = 0; k < ns; k++) {
                if (spec[k] == -1) { if (infer >= 0) return -1; infer = (int)k; }
                else if (spec[k] < 0) return -1;
                else { r.s.d[k] = (size_t)spec[k]; known *= r.s.d[k]; }
            }
            if (infer >= 0) { if (known == 0 || total % known) return -1; r.s.d[infer] = total / known; }
            if (shp_count(r.s) != total) return -1;
            r.dt = x.dt; r.data = x.data;
            break;
        }
        case K_EXPAND: {
            val a;
            if (get_in(vals, op, 0, &x) || get_in(vals, op, 1, &a) || a.dt != DT_I32 || shp_count(a.s) != 1) return -1;
            int32_t axis = ((const int32_t *)a.data)[0];
            int rr = x.s.rank;
            int ax = axis < 0 ? axis + rr + 1 : axis;
            if (ax < 0 || ax > rr || rr >= MAXR) return -1;
            r.s.rank = rr + 1;
            for (int k = 0; k < ax; k++) r.s.d[k] = x.s.d[k];
            r.s.d[ax] = 1;
            for (int k = ax; k < rr; k++) r.s.d[k + 1] = x.s.d[k];
            r.dt = x.dt; r.data = x.data;
            break;
        }
        case K_SHAPE: {
            if (get_in(vals, op, 0, &x)) return -1;
            int32_t *o = (int32_t *)sc_alloc(sc, 4 * (size_t)x.s.rank);
            for (int k = 0; k < x.s.rank; k++) o[k] = (int32_t)x.s.d[k];
            r.s.rank = 1; r.s.d[0] = (size_t)x.s.rank; r.dt = DT_I32; r.data = o;
            break;
        }
        case K_FILL: {
            val dv, v;
            if (get_in(vals, op, 0, &dv) || get_in(vals, op, 1, &v) || dv.dt != DT_I32) return -1;
            size_t nd = shp_count(dv.s);
            if (nd > MAXR) return -1;
            r.s.rank = (int)nd;
            for (size_t k = 0; k < nd; k++) {
                int32_t d = ((const int32_t *)dv.data)[k];
                if (d < 0) return -1;
                r.s.d[k] = (size_t)d;
            }
            if (shp_count(v.s) != 1) return -1;
            size_t n = shp_count(r.s), es = v.dt == DT_BOOL ? 1 : 4;
            uint8_t *o = (uint8_t *)sc_alloc(sc, es * n);
            for (size_t i = 0; i < n; i++) memcpy(o + i * es, v.data, es);
            r.dt = v.dt; r.data = o;
            break;
        }
        case K_SPLIT: {
            val a;
            if (get_in(vals, op, 0, &a) || a.dt != DT_I32 || shp_count(a.s) != 1) return -1;
            if (get_in(vals, op, 1, &x) || x.dt != DT_F32) return -1;
            int ax;
            if (norm_axis(((const int32_t *)a.data)[0], x.s.rank, &ax)) return -1;
            size_t parts = op->nout;
            if (parts == 0 || x.s.d[ax] % parts) return -1;
            shp s = x.s;
            s.d[ax] /= parts;
            size_t outer = 1, inner = 1;
            for (int k = 0; k < ax; k++) outer *= x.s.d[k];
            for (int k = ax + 1; k < x.s.rank; k++) inner *= x.s.d[k];
            size_t chunk = s.d[ax] * inner;
            for (size_t p = 0; p < parts; p++) {
                float *o = (float *)sc_alloc(sc, 4 * outer * chunk);
                for (size_t oi = 0; oi < outer; oi++)
                    memcpy(o + oi * chunk, (const float *)x.data + (oi * parts + p) * chunk, 4 * chunk);
                val pv = {1, s, DT_F32, o};
                vals[op->out[p]] = pv;
            }
            return 0;
        }
        default: return -1;
    }
    vals[out0] = r;
    return 0;
}

/* lattice.zig:49-71 (Mapper.predict) over tflite.zig:378-388 (Model.run) */
int vo_lattice_eval(const vo_lattice *m, const double *f0, const double *f1, const double *f2, const double *f3,
                    double *y) {
    scratch sc = {0};
    val *vals = (val *)malloc(sizeof(val) * m->ntens);
    memcpy(vals, m->consts, sizeof(val) * m->ntens);
    const double *feats[4] = {f0, f1, f2, f3};
    size_t nb = m->nb;
    float *scal = (float *)sc_alloc(&sc, sizeof(float) * (4 * nb + 1));
    for (int fi = 0; fi < 4; fi++)
        for (size_t b = 0; b < nb; b++) {
            size_t k = fi * nb + b;
            scal[k] = (float)feats[fi][b];
            val v = {1, {1, {1}}, DT_F32, &scal[k]};
            vals[m->feat[k]] = v;
        }
    scal[4 * nb] = 0.5f;
    val tv = {1, {1, {1}}, DT_F32, &scal[4 * nb]};
    vals[m->tau] = tv;
    int rc = VO_OK;
    for (size_t i = 0; i < m->nops; i++)
        if (exec_op(m, &sc, &m->ops[i], vals)) { rc = VO_E_MODEL; break; }
    if (rc == VO_OK) {
        val o = vals[m->output];
        if (!o.set || o.dt != DT_F32 || shp_count(o.s) < 1) rc = VO_E_MODEL;
        else *y = (double)((const float *)o.data)[0];
    }
    free(vals);
    sc_free(&sc);
    return rc;
}

int vo_lattice_predict(const uint8_t *bytes, size_t len, const double *a, const double *b, const double *c,
                       const double *d, size_t nb, double *y) {
    int err;
    vo_lattice *m = vo_lattice_load(bytes, len, nb, &err);
    if (!m) return err;
    int rc = vo_lattice_eval(m, a, b, c, d, y);
    vo_lattice_free(m);
    return rc;
}

int vo_lattice_op_count(const uint8_t *bytes, size_t len) {
    int err;
    vo_lattice *m = vo_lattice_load(bytes, len, 21, &err);
    if (!m) return err;
    int n = (int)m->nops;
    vo_lattice_free(m);
    return n;
}
