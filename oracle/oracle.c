/*
 * oracle.c -- CPU restatement of the Beamcast Go renderer's hot path (test infrastructure).
 * Every function cites the windows_portable/ source it restates.  See oracle.h.
 * Compile with -ffp-contract=off: the fp32 paths must evaluate exactly the operations
 * written, in the order written, so that they can be compared bit for bit with the GPU.
 */
#include "oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* ------------------------------------------------------------------ fp64 vectors
 * renderer.go:21-47 (Vec3 ops, evaluation order of Dot/Cross kept). */
typedef struct { double x, y, z; } v3;
static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double len2(v3 a) { return dot(a, a); }
static inline v3 unit(v3 a) {
    double l = sqrt(len2(a));
    if (l < 1e-12) return V(0, 0, 0);
    return mul(a, 1 / l);
}

typedef struct { v3 mn, mx; } box3;
static box3 empty_box(void) {
    box3 b = {{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    return b;
}
/* renderer.go:117-127 ExpandPoint/ExpandBox (math.Min/Max; no NaNs occur here) */
static inline box3 expand_pt(box3 b, v3 p) {
    if (p.x < b.mn.x) b.mn.x = p.x;
    if (p.y < b.mn.y) b.mn.y = p.y;
    if (p.z < b.mn.z) b.mn.z = p.z;
    if (p.x > b.mx.x) b.mx.x = p.x;
    if (p.y > b.mx.y) b.mx.y = p.y;
    if (p.z > b.mx.z) b.mx.z = p.z;
    return b;
}
static inline box3 expand_box(box3 b, box3 c) { return expand_pt(expand_pt(b, c.mn), c.mx); }
/* renderer.go:129-135 */
static inline double surface_area(box3 b) {
    v3 e = sub(b.mx, b.mn);
    if (e.x < 0 || e.y < 0 || e.z < 0) return 0;
    return 2 * (e.x * e.y + e.y * e.z + e.z * e.x);
}

typedef struct {
    int kind, mat;
    v3 center;
    double radius;
    v3 a, b, c, e1, e2, n;
    box3 box;
    v3 centroid;
} prim_t;
typedef struct { int kind; v3 color; double rough, ior; v3 emission; } mat_t;
typedef struct { box3 box; int left, right, start, count; } node_t;

/* fp32 records: what the reference packer uploads (gpu_opencl_windows.go:337-377), with node
 * boxes that cover the rounded records and are rounded outward, so that culling stays
 * conservative for the fp32 primitive tests. */
typedef struct { float mn[3], mx[3]; int escape; uint32_t meta; } node32_t;
typedef struct {
    int kind, mat, id;
    float a[3], e1[3], e2[3];         /* triangle */
    float c[3], r, r2, invr;          /* sphere   */
} prim32_t;
typedef struct { int kind; float color[3], rough, ior, emission[3]; } mat32_t;

typedef struct {
    int prim;          /* Scene.Primitives index */
    int kind;          /* sphere / triangle */
    double area;       /* surface area in fp64 */
} light_t;

struct orc_scene {
    int nmat, nprim, nnode, cap;
    mat_t* mats;
    prim_t* prims;
    node_t* nodes;
    int* order;
    int bins, leaf, nee_tri, fallback_sorts;
    int nlight;
    light_t* lights;
    double* cdf;
    double* pdf_by_prim;
    int* escape;
    node32_t* n32;
    prim32_t* p32;      /* indexed by Scene.Primitives index */
    mat32_t* m32;
    float* cdf32;
    float* pdf32;       /* by primitive */
};

/* ------------------------------------------------------------------ scene build */
static double luminance(v3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; } /* renderer.go:2142 */

static double centroid_axis(v3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }
/* renderer.go:486-494 */
static int largest_axis(v3 e) {
    if (e.y > e.x && e.y >= e.z) return 1;
    if (e.z > e.x && e.z > e.y) return 2;
    return 0;
}

static int push_node(orc_scene* s) {
    if (s->nnode == s->cap) {
        s->cap = s->cap ? s->cap * 2 : 64;
        s->nodes = (node_t*)realloc(s->nodes, sizeof(node_t) * (size_t)s->cap);
    }
    node_t z = {empty_box(), -1, -1, 0, 0};
    s->nodes[s->nnode] = z;
    return s->nnode++;
}

static int g_sort_axis;
static const orc_scene* g_sort_scene;
static int cmp_centroid(const void* a, const void* b) {
    int ia = *(const int*)a, ib = *(const int*)b;
    double ca = centroid_axis(g_sort_scene->prims[ia].centroid, g_sort_axis);
    double cb = centroid_axis(g_sort_scene->prims[ib].centroid, g_sort_axis);
    if (ca < cb) return -1;
    if (ca > cb) return 1;
    return ia < ib ? -1 : (ia > ib);
}

static inline int bin_of(double c, double minC, double invExtent, int binsN) {
    int b = (int)((c - minC) * invExtent * (double)binsN);
    if (b < 0) b = 0;
    if (b >= binsN) b = binsN - 1;
    return b;
}

/* buildNodeSAH, renderer.go:526-633 */
static int build_sah(orc_scene* s, int start, int end) {
    int idx = push_node(s);
    box3 box = empty_box(), cb = empty_box();
    for (int i = start; i < end; i++) { /* nodeBounds, renderer.go:496-504 */
        const prim_t* p = &s->prims[s->order[i]];
        box = expand_box(box, p->box);
        cb = expand_pt(cb, p->centroid);
    }
    int count = end - start;
    node_t leafn = {box, -1, -1, start, count};
    if (count <= s->leaf) { s->nodes[idx] = leafn; return idx; }
    v3 ext = sub(cb.mx, cb.mn);
    int axis = largest_axis(ext);
    double axisExtent = centroid_axis(cb.mx, axis) - centroid_axis(cb.mn, axis);
    if (axisExtent < 1e-12) { s->nodes[idx] = leafn; return idx; }
    int binsN = s->bins;
    box3 bbox[32];
    int bcount[32];
    for (int i = 0; i < binsN; i++) { bbox[i] = empty_box(); bcount[i] = 0; }
    double minC = centroid_axis(cb.mn, axis);
    double invExtent = 1 / axisExtent;
    for (int i = start; i < end; i++) {
        const prim_t* p = &s->prims[s->order[i]];
        int b = bin_of(centroid_axis(p->centroid, axis), minC, invExtent, binsN);
        bcount[b]++;
        bbox[b] = expand_box(bbox[b], p->box);
    }
    int leftCount[31], rightCount[31];
    double leftArea[31], rightArea[31];
    box3 acc = empty_box();
    int accCount = 0;
    for (int i = 0; i < binsN - 1; i++) {
        if (bcount[i] > 0) acc = expand_box(acc, bbox[i]);
        accCount += bcount[i];
        leftCount[i] = accCount;
        leftArea[i] = surface_area(acc);
    }
    acc = empty_box();
    accCount = 0;
    for (int i = binsN - 1; i > 0; i--) {
        if (bcount[i] > 0) acc = expand_box(acc, bbox[i]);
        accCount += bcount[i];
        rightCount[i - 1] = accCount;
        rightArea[i - 1] = surface_area(acc);
    }
    int bestSplit = -1;
    double bestCost = INFINITY;
    for (int i = 0; i < binsN - 1; i++) {
        if (leftCount[i] == 0 || rightCount[i] == 0) continue;
        double cost = leftArea[i] * (double)leftCount[i] + rightArea[i] * (double)rightCount[i];
        if (cost < bestCost) { bestCost = cost; bestSplit = i; }
    }
    double leafCost = surface_area(box) * (double)count;
    if (bestSplit < 0 || (count <= s->leaf * 2 && bestCost >= leafCost * 0.98)) {
        s->nodes[idx] = leafn;
        return idx;
    }
    int i = start, j = end - 1;
    while (i <= j) {
        const prim_t* p = &s->prims[s->order[i]];
        int b = bin_of(centroid_axis(p->centroid, axis), minC, invExtent, binsN);
        if (b <= bestSplit) {
            i++;
        } else {
            int t = s->order[i]; s->order[i] = s->order[j]; s->order[j] = t;
            j--;
        }
    }
    int mid = i;
    if (mid <= start || mid >= end) { /* renderer.go:623-628; unreachable for SAH splits */
        s->fallback_sorts++;
        g_sort_axis = axis;
        g_sort_scene = s;
        qsort(s->order + start, (size_t)count, sizeof(int), cmp_centroid);
        mid = (start + end) / 2;
    }
    int left = build_sah(s, start, mid);
    int right = build_sah(s, mid, end);
    node_t in = {box, left, right, 0, 0};
    s->nodes[idx] = in;
    return idx;
}

/* threadedBVHEscapes, renderer.go:3251-3276 */
static void assign_escape(orc_scene* s, int idx, int next) {
    while (idx >= 0 && idx < s->nnode) {
        s->escape[idx] = next;
        const node_t* n = &s->nodes[idx];
        if (n->count > 0) return;
        if (!(n->left >= 0 && n->right >= 0)) return;
        assign_escape(s, n->left, n->right);
        idx = n->right; /* tail position: escape of right child = parent's escape */
    }
}

/* lightWeight (renderer.go:1343-1357) extended to emissive triangles when requested. */
static void rebuild_lights(orc_scene* s) { /* rebuildLightDistribution, renderer.go:363-407 */
    s->lights = (light_t*)calloc((size_t)s->nprim + 1, sizeof(light_t));
    s->nlight = 0;
    for (int i = 0; i < s->nprim; i++) {
        const prim_t* p = &s->prims[i];
        if (p->mat < 0 || p->mat >= s->nmat || s->mats[p->mat].kind != ORC_EMISSIVE) continue;
        if (p->kind == ORC_SPHERE) {
            light_t L = {i, ORC_SPHERE, 4 * M_PI * p->radius * p->radius};
            s->lights[s->nlight++] = L;
        } else if (s->nee_tri) {
            light_t L = {i, ORC_TRIANGLE, 0.5 * sqrt(len2(cross(p->e1, p->e2)))};
            s->lights[s->nlight++] = L;
        }
    }
    s->cdf = (double*)calloc((size_t)s->nlight + 1, sizeof(double));
    s->pdf_by_prim = (double*)calloc((size_t)s->nprim + 1, sizeof(double));
    double total = 0;
    for (int i = 0; i < s->nlight; i++) {
        const light_t* L = &s->lights[i];
        double w = luminance(s->mats[s->prims[L->prim].mat].emission) * L->area;
        if (w < 1e-9) w = 1e-9;
        total += w;
        s->cdf[i] = total;
    }
    if (total <= 0) {
        if (s->nlight == 0) return;
        double pdf = 1 / (double)s->nlight;
        for (int i = 0; i < s->nlight; i++) {
            s->cdf[i] = (double)(i + 1) / (double)s->nlight;
            s->pdf_by_prim[s->lights[i].prim] = pdf;
        }
        return;
    }
    double prev = 0;
    for (int i = 0; i < s->nlight; i++) {
        double cdf = s->cdf[i] / total;
        s->cdf[i] = cdf;
        s->pdf_by_prim[s->lights[i].prim] = cdf - prev;
        prev = cdf;
    }
    if (s->nlight > 0) s->cdf[s->nlight - 1] = 1;
}

static float f_down(double v) { float f = (float)v; if ((double)f > v) f = nextafterf(f, -INFINITY); return f; }
static float f_up(double v) { float f = (float)v; if ((double)f < v) f = nextafterf(f, INFINITY); return f; }

orc_scene* orc_scene_new(int nmat, const int32_t* mat_kind, const double* md, int nprim,
                         const int32_t* prim_kind, const int32_t* prim_mat, const double* pd,
                         int bins, int leaf, int nee_triangles) {
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->nmat = nmat;
    s->mats = (mat_t*)calloc((size_t)nmat + 1, sizeof(mat_t));
    for (int i = 0; i < nmat; i++) {
        const double* m = md + 8 * (size_t)i;
        mat_t mm = {mat_kind[i], V(m[0], m[1], m[2]), m[3], m[4], V(m[5], m[6], m[7])};
        s->mats[i] = mm;
    }
    s->prims = (prim_t*)calloc((size_t)nprim + 1, sizeof(prim_t));
    for (int i = 0; i < nprim; i++) {
        const double* g = pd + 9 * (size_t)i;
        prim_t p;
        memset(&p, 0, sizeof p);
        p.mat = prim_mat[i];
        if (prim_kind[i] == ORC_SPHERE) { /* AddSphere, renderer.go:246-256 */
            double r = g[3];
            if (r <= 0) continue;
            v3 c = V(g[0], g[1], g[2]), rv = V(r, r, r);
            p.kind = ORC_SPHERE;
            p.center = c;
            p.radius = r;
            p.centroid = c;
            p.box.mn = sub(c, rv);
            p.box.mx = add(c, rv);
        } else { /* AddTriangle, renderer.go:257-270 */
            v3 a = V(g[0], g[1], g[2]), b = V(g[3], g[4], g[5]), c = V(g[6], g[7], g[8]);
            v3 e1 = sub(b, a), e2 = sub(c, a), n = cross(e1, e2);
            if (len2(n) < 1e-18) continue;
            box3 bx = expand_pt(expand_pt(expand_pt(empty_box(), a), b), c);
            double eps = 1e-6;
            bx.mn = sub(bx.mn, V(eps, eps, eps));
            bx.mx = add(bx.mx, V(eps, eps, eps));
            p.kind = ORC_TRIANGLE;
            p.a = a; p.b = b; p.c = c; p.e1 = e1; p.e2 = e2; p.n = unit(n);
            p.box = bx;
            p.centroid = mul(add(add(a, b), c), 1.0 / 3.0);
        }
        s->prims[s->nprim++] = p;
    }
    /* Scene.Build parameter clamps, renderer.go:409-427 */
    if (bins < 4) bins = 16;
    if (bins > 32) bins = 32;
    if (leaf < 1) leaf = 4;
    if (leaf > 16) leaf = 16;
    s->bins = bins;
    s->leaf = leaf;
    s->nee_tri = nee_triangles;
    rebuild_lights(s);
    s->order = (int*)malloc(sizeof(int) * ((size_t)s->nprim + 1));
    for (int i = 0; i < s->nprim; i++) s->order[i] = i;
    if (s->nprim > 0) build_sah(s, 0, s->nprim);
    s->escape = (int*)malloc(sizeof(int) * ((size_t)s->nnode + 1));
    for (int i = 0; i < s->nnode; i++) s->escape[i] = -1;
    if (s->nnode > 0) assign_escape(s, 0, -1);

    /* fp32 records (gpu_opencl_windows.go:337-377) */
    s->p32 = (prim32_t*)calloc((size_t)s->nprim + 1, sizeof(prim32_t));
    for (int i = 0; i < s->nprim; i++) {
        const prim_t* p = &s->prims[i];
        prim32_t q;
        memset(&q, 0, sizeof q);
        q.kind = p->kind; q.mat = p->mat; q.id = i;
        if (p->kind == ORC_TRIANGLE) {
            q.a[0] = (float)p->a.x; q.a[1] = (float)p->a.y; q.a[2] = (float)p->a.z;
            q.e1[0] = (float)p->e1.x; q.e1[1] = (float)p->e1.y; q.e1[2] = (float)p->e1.z;
            q.e2[0] = (float)p->e2.x; q.e2[1] = (float)p->e2.y; q.e2[2] = (float)p->e2.z;
        } else {
            q.c[0] = (float)p->center.x; q.c[1] = (float)p->center.y; q.c[2] = (float)p->center.z;
            q.r = (float)p->radius;
            q.r2 = (float)(p->radius * p->radius);
            q.invr = (float)(1.0 / p->radius);
        }
        s->p32[i] = q;
    }
    /* Node boxes for the fp32 traversal.  The fp32 primitive test sees the rounded record
     * (a triangle a + u e1 + v e2 of fp32 a, e1, e2; a sphere of fp32 centre and r^2), which
     * can lie outside the fp64 box by half an ulp of the coordinates -- far from the origin a
     * large fraction of a small triangle.  Each box therefore covers the fp64 box and the
     * extent of the rounded records (a + e1, a + e2 are exact in double), is widened by
     * 2^-21 (|coordinate| + extent) for the rounding inside the box and primitive tests, and
     * is rounded outward to fp32: culling never drops a ray that the fp32 test would hit, so
     * the result equals a test of every primitive.  Children come after their parent in
     * s->nodes, so one reverse pass builds the boxes bottom-up. */
    s->n32 = (node32_t*)calloc((size_t)s->nnode + 1, sizeof(node32_t));
    box3* eb = (box3*)malloc(sizeof(box3) * ((size_t)s->nnode + 1));
    for (int i = s->nnode - 1; i >= 0; i--) {
        const node_t* n = &s->nodes[i];
        box3 b = n->box;
        if (n->count > 0) {
            for (int k = n->start; k < n->start + n->count; k++) {
                const prim32_t* q = &s->p32[s->order[k]];
                if (q->kind == ORC_TRIANGLE) {
                    v3 a = V(q->a[0], q->a[1], q->a[2]);
                    b = expand_pt(b, a);
                    b = expand_pt(b, add(a, V(q->e1[0], q->e1[1], q->e1[2])));
                    b = expand_pt(b, add(a, V(q->e2[0], q->e2[1], q->e2[2])));
                } else {
                    double r = sqrt((double)q->r2);
                    b = expand_pt(b, V(q->c[0] - r, q->c[1] - r, q->c[2] - r));
                    b = expand_pt(b, V(q->c[0] + r, q->c[1] + r, q->c[2] + r));
                }
            }
        } else {
            b = expand_box(b, eb[n->left]);
            b = expand_box(b, eb[n->right]);
        }
        eb[i] = b;
        double lo[3] = {b.mn.x, b.mn.y, b.mn.z}, hi[3] = {b.mx.x, b.mx.y, b.mx.z};
        node32_t o;
        for (int k = 0; k < 3; k++) {
            const double pad = ldexp(fmax(fabs(lo[k]), fabs(hi[k])) + (hi[k] - lo[k]), -21);
            o.mn[k] = f_down(lo[k] - pad);
            o.mx[k] = f_up(hi[k] + pad);
        }
        o.escape = s->escape[i];
        if (n->count > 0)
            o.meta = 0x80000000u | ((uint32_t)(n->count - 1) & 0xf) << 27 | ((uint32_t)n->start & 0x07ffffffu);
        else
            o.meta = (uint32_t)n->left & 0x07ffffffu;
        s->n32[i] = o;
    }
    free(eb);
    s->m32 = (mat32_t*)calloc((size_t)s->nmat + 1, sizeof(mat32_t));
    for (int i = 0; i < s->nmat; i++) {
        const mat_t* m = &s->mats[i];
        mat32_t q = {m->kind, {(float)m->color.x, (float)m->color.y, (float)m->color.z}, (float)m->rough,
                     (float)m->ior, {(float)m->emission.x, (float)m->emission.y, (float)m->emission.z}};
        s->m32[i] = q;
    }
    s->cdf32 = (float*)calloc((size_t)s->nlight + 1, sizeof(float));
    for (int i = 0; i < s->nlight; i++) s->cdf32[i] = (float)s->cdf[i];
    s->pdf32 = (float*)calloc((size_t)s->nprim + 1, sizeof(float));
    for (int i = 0; i < s->nprim; i++) s->pdf32[i] = (float)s->pdf_by_prim[i];
    return s;
}

void orc_scene_free(orc_scene* s) {
    if (!s) return;
    free(s->mats); free(s->prims); free(s->nodes); free(s->order); free(s->lights); free(s->cdf);
    free(s->pdf_by_prim); free(s->escape); free(s->n32); free(s->p32); free(s->m32); free(s->cdf32);
    free(s->pdf32); free(s);
}
int orc_scene_nprim(const orc_scene* s) { return s->nprim; }
int orc_scene_nnode(const orc_scene* s) { return s->nnode; }
int orc_scene_fallback_sorts(const orc_scene* s) { return s->fallback_sorts; }
void orc_scene_nodes(const orc_scene* s, double* box, int32_t* link, int32_t* order) {
    for (int i = 0; i < s->nnode; i++) {
        const node_t* n = &s->nodes[i];
        double* b = box + 6 * (size_t)i;
        b[0] = n->box.mn.x; b[1] = n->box.mn.y; b[2] = n->box.mn.z;
        b[3] = n->box.mx.x; b[4] = n->box.mx.y; b[5] = n->box.mx.z;
        int32_t* l = link + 4 * (size_t)i;
        l[0] = n->left; l[1] = n->right; l[2] = n->start; l[3] = n->count;
    }
    for (int i = 0; i < s->nprim; i++) order[i] = s->order[i];
}
void orc_scene_escapes(const orc_scene* s, int32_t* e) { for (int i = 0; i < s->nnode; i++) e[i] = s->escape[i]; }
int orc_scene_nlights(const orc_scene* s) { return s->nlight; }
void orc_scene_lights(const orc_scene* s, int32_t* lp, double* cdf, double* pdf) {
    for (int i = 0; i < s->nlight; i++) { lp[i] = s->lights[i].prim; cdf[i] = s->cdf[i]; }
    for (int i = 0; i < s->nprim; i++) pdf[i] = s->pdf_by_prim[i];
}

/* ------------------------------------------------------------------ fp64 traversal (Go semantics) */
typedef struct {
    v3 o, d, inv;
    double dirlen2, invdirlen2;
    int zx, zy, zz;
} rayacc;
/* prepareRay, renderer.go:52-77 */
static rayacc prepare(v3 o, v3 d) {
    rayacc a;
    memset(&a, 0, sizeof a);
    a.o = o; a.d = d;
    a.dirlen2 = len2(d);
    if (a.dirlen2 > 1e-18) a.invdirlen2 = 1 / a.dirlen2;
    if (fabs(d.x) < 1e-15) a.zx = 1; else a.inv.x = 1 / d.x;
    if (fabs(d.y) < 1e-15) a.zy = 1; else a.inv.y = 1 / d.y;
    if (fabs(d.z) < 1e-15) a.zz = 1; else a.inv.z = 1 / d.z;
    return a;
}
/* AABB.HitAccel, renderer.go:147-210 */
static int box_hit(const box3* b, const rayacc* r, double tmin, double tmax, double* nearp) {
    double near = tmin;
    const double o[3] = {r->o.x, r->o.y, r->o.z}, inv[3] = {r->inv.x, r->inv.y, r->inv.z};
    const double mn[3] = {b->mn.x, b->mn.y, b->mn.z}, mx[3] = {b->mx.x, b->mx.y, b->mx.z};
    const int z[3] = {r->zx, r->zy, r->zz};
    for (int k = 0; k < 3; k++) {
        if (z[k]) {
            if (o[k] < mn[k] || o[k] > mx[k]) return 0;
        } else {
            double t0 = (mn[k] - o[k]) * inv[k], t1 = (mx[k] - o[k]) * inv[k];
            if (inv[k] < 0) { double t = t0; t0 = t1; t1 = t; }
            if (t0 > near) near = t0;
            if (t1 < tmax) tmax = t1;
            if (tmax <= near) return 0;
        }
    }
    if (nearp) *nearp = near;
    return 1;
}
int orc_aabb_hit(const double* b6, const double* r6, double tmin, double tmax, double* near_out) {
    box3 b = {{b6[0], b6[1], b6[2]}, {b6[3], b6[4], b6[5]}};
    rayacc r = prepare(V(r6[0], r6[1], r6[2]), V(r6[3], r6[4], r6[5]));
    double nr = 0;
    int h = box_hit(&b, &r, tmin, tmax, &nr);
    if (near_out) *near_out = h ? nr : 0;
    return h;
}

typedef struct { double t; v3 n; int mat, prim, front; } hit64;

/* sphereHitAccel, renderer.go:650-678 */
static int sphere_hit(const prim_t* p, const rayacc* ar, double tmin, double tmax, hit64* h) {
    if (ar->invdirlen2 <= 0) return 0;
    v3 oc = sub(ar->o, p->center);
    double hb = dot(oc, ar->d);
    double c = len2(oc) - p->radius * p->radius;
    double disc = hb * hb - ar->dirlen2 * c;
    if (disc < 0) return 0;
    double sq = sqrt(disc);
    double root = (-hb - sq) * ar->invdirlen2;
    if (root < tmin || root > tmax) {
        root = (-hb + sq) * ar->invdirlen2;
        if (root < tmin || root > tmax) return 0;
    }
    v3 pos = add(ar->o, mul(ar->d, root));
    v3 outward = mul(sub(pos, p->center), 1 / p->radius);
    int front = dot(ar->d, outward) < 0;
    h->t = root; h->n = front ? outward : mul(outward, -1); h->mat = p->mat; h->front = front;
    return 1;
}
/* triHit, renderer.go:684-711 */
static int tri_hit(const prim_t* p, v3 o, v3 d, double tmin, double tmax, hit64* hh) {
    v3 h = cross(d, p->e2);
    double det = dot(p->e1, h);
    if (fabs(det) < 1e-10) return 0;
    double f = 1 / det;
    v3 ss = sub(o, p->a);
    double u = f * dot(ss, h);
    if (u < 0 || u > 1) return 0;
    v3 q = cross(ss, p->e1);
    double v = f * dot(d, q);
    if (v < 0 || u + v > 1) return 0;
    double t = f * dot(p->e2, q);
    if (t < tmin || t > tmax) return 0;
    v3 n = p->n;
    int front = dot(d, n) < 0;
    hh->t = t; hh->n = front ? n : mul(n, -1); hh->mat = p->mat; hh->front = front;
    return 1;
}
static int prim_hit(const prim_t* p, const rayacc* ar, double tmin, double tmax, hit64* h) {
    return p->kind == ORC_SPHERE ? sphere_hit(p, ar, tmin, tmax, h) : tri_hit(p, ar->o, ar->d, tmin, tmax, h);
}
/* sphereAnyHitAccel / triAnyHit, renderer.go:724-770 */
static int prim_any(const prim_t* p, const rayacc* ar, double tmin, double tmax) {
    hit64 h;
    return prim_hit(p, ar, tmin, tmax, &h);
}

static int hit_rec(const orc_scene* s, int ni, const rayacc* ar, double tmin, double tmax, hit64* best) {
    /* hitNodeRecursiveAccel, renderer.go:917-939 */
    const node_t* n = &s->nodes[ni];
    if (!box_hit(&n->box, ar, tmin, tmax, NULL)) return 0;
    double closest = tmax;
    int found = 0;
    if (n->count > 0) {
        for (int i = 0; i < n->count; i++) {
            int pi = s->order[n->start + i];
            hit64 h;
            if (prim_hit(&s->prims[pi], ar, tmin, closest, &h)) {
                h.prim = pi; closest = h.t; *best = h; found = 1;
            }
        }
        return found;
    }
    hit64 h;
    if (hit_rec(s, n->left, ar, tmin, closest, &h)) { closest = h.t; *best = h; found = 1; }
    if (hit_rec(s, n->right, ar, tmin, closest, &h)) { *best = h; found = 1; }
    return found;
}

/* Scene.Hit, renderer.go:777-849 */
static int scene_hit64(const orc_scene* s, v3 o, v3 d, double tmin, double tmax, hit64* best) {
    if (s->nnode == 0) return 0;
    rayacc ar = prepare(o, d);
    double rootNear;
    if (!box_hit(&s->nodes[0].box, &ar, tmin, tmax, &rootNear)) return 0;
    int stn[16];
    double stt[16];
    int sp = 1;
    stn[0] = 0; stt[0] = rootNear;
    double closest = tmax;
    int found = 0;
    while (sp > 0) {
        sp--;
        int ni = stn[sp];
        if (stt[sp] >= closest) continue;
        const node_t* n = &s->nodes[ni];
        if (n->count > 0) {
            for (int i = 0; i < n->count; i++) {
                int pi = s->order[n->start + i];
                hit64 h;
                if (prim_hit(&s->prims[pi], &ar, tmin, closest, &h)) {
                    h.prim = pi; closest = h.t; *best = h; found = 1;
                }
            }
            continue;
        }
        double ln = 0, rn = 0;
        int lok = box_hit(&s->nodes[n->left].box, &ar, tmin, closest, &ln);
        int rok = box_hit(&s->nodes[n->right].box, &ar, tmin, closest, &rn);
        if (!lok && !rok) continue;
        if (sp + 2 > 16) {
            hit64 h;
            if (lok && hit_rec(s, n->left, &ar, tmin, closest, &h)) { closest = h.t; *best = h; found = 1; }
            if (rok && hit_rec(s, n->right, &ar, tmin, closest, &h)) { closest = h.t; *best = h; found = 1; }
            continue;
        }
        if (lok && rok) {
            if (ln <= rn) {
                stn[sp] = n->right; stt[sp] = rn; sp++;
                stn[sp] = n->left; stt[sp] = ln; sp++;
            } else {
                stn[sp] = n->left; stt[sp] = ln; sp++;
                stn[sp] = n->right; stt[sp] = rn; sp++;
            }
        } else if (lok) {
            stn[sp] = n->left; stt[sp] = ln; sp++;
        } else {
            stn[sp] = n->right; stt[sp] = rn; sp++;
        }
    }
    return found;
}

static int occ_rec(const orc_scene* s, int ni, const rayacc* ar, double tmin, double tmax) {
    /* occludedNodeRecursive, renderer.go:897-915 */
    const node_t* n = &s->nodes[ni];
    if (!box_hit(&n->box, ar, tmin, tmax, NULL)) return 0;
    if (n->count > 0) {
        for (int i = 0; i < n->count; i++)
            if (prim_any(&s->prims[s->order[n->start + i]], ar, tmin, tmax)) return 1;
        return 0;
    }
    return occ_rec(s, n->left, ar, tmin, tmax) || occ_rec(s, n->right, ar, tmin, tmax);
}
/* Scene.Occluded, renderer.go:851-895 */
static int scene_occluded64(const orc_scene* s, v3 o, v3 d, double tmin, double tmax) {
    if (s->nnode == 0) return 0;
    rayacc ar = prepare(o, d);
    if (!box_hit(&s->nodes[0].box, &ar, tmin, tmax, NULL)) return 0;
    int st[16];
    int sp = 1;
    st[0] = 0;
    while (sp > 0) {
        sp--;
        const node_t* n = &s->nodes[st[sp]];
        if (n->count > 0) {
            for (int i = 0; i < n->count; i++)
                if (prim_any(&s->prims[s->order[n->start + i]], &ar, tmin, tmax)) return 1;
            continue;
        }
        int lok = box_hit(&s->nodes[n->left].box, &ar, tmin, tmax, NULL);
        int rok = box_hit(&s->nodes[n->right].box, &ar, tmin, tmax, NULL);
        if (sp + 2 > 16) {
            if (lok && occ_rec(s, n->left, &ar, tmin, tmax)) return 1;
            if (rok && occ_rec(s, n->right, &ar, tmin, tmax)) return 1;
            continue;
        }
        if (lok) st[sp++] = n->left;
        if (rok) st[sp++] = n->right;
    }
    return 0;
}

void orc_hit64(const orc_scene* s, int64_t n, const double* rays, int32_t* prim, double* t) {
    for (int64_t i = 0; i < n; i++) {
        const double* r = rays + 8 * i;
        hit64 h;
        if (scene_hit64(s, V(r[0], r[1], r[2]), V(r[4], r[5], r[6]), r[3], r[7], &h)) {
            prim[i] = h.prim; t[i] = h.t;
        } else {
            prim[i] = -1; t[i] = 0;
        }
    }
}
/* the threaded closest-hit of TestThreadedBVHClosestHitEquivalence, renderer_test.go:797-830 */
void orc_hit64_threaded(const orc_scene* s, int64_t n, const double* rays, int32_t* prim, double* t) {
    for (int64_t k = 0; k < n; k++) {
        const double* r = rays + 8 * k;
        rayacc ar = prepare(V(r[0], r[1], r[2]), V(r[4], r[5], r[6]));
        double tmin = r[3], bestT = r[7];
        int found = 0, bp = -1;
        for (int ni = 0; ni >= 0 && ni < s->nnode;) {
            const node_t* nd = &s->nodes[ni];
            if (!box_hit(&nd->box, &ar, tmin, bestT, NULL)) { ni = s->escape[ni]; continue; }
            if (nd->count > 0) {
                for (int j = 0; j < nd->count; j++) {
                    int pi = s->order[nd->start + j];
                    hit64 h;
                    if (prim_hit(&s->prims[pi], &ar, tmin, bestT, &h)) { bestT = h.t; bp = pi; found = 1; }
                }
                ni = s->escape[ni];
            } else {
                ni = nd->left;
            }
        }
        prim[k] = found ? bp : -1;
        t[k] = found ? bestT : 0;
    }
}
void orc_occluded64(const orc_scene* s, int64_t n, const double* rays, int32_t* occ) {
    for (int64_t i = 0; i < n; i++) {
        const double* r = rays + 8 * i;
        occ[i] = scene_occluded64(s, V(r[0], r[1], r[2]), V(r[4], r[5], r[6]), r[3], r[7]);
    }
}
void orc_brute64(const orc_scene* s, int64_t n, const double* rays, int32_t* prim, double* t) {
    for (int64_t k = 0; k < n; k++) {
        const double* r = rays + 8 * k;
        rayacc ar = prepare(V(r[0], r[1], r[2]), V(r[4], r[5], r[6]));
        double best = r[7];
        int bp = -1;
        for (int i = 0; i < s->nprim; i++) {
            hit64 h;
            if (prim_hit(&s->prims[i], &ar, r[3], best, &h)) { best = h.t; bp = i; }
        }
        prim[k] = bp;
        t[k] = bp >= 0 ? best : 0;
    }
}

/* ------------------------------------------------------------------ fp32 traversal
 * Same triangle arithmetic as triHit (renderer.go:684-711) in single precision; sphere
 * uses the line-distance discriminant (Haines et al., "Precision Improvements for
 * Ray/Sphere Intersection", Ray Tracing Gems 2019, ch. 7) because the reference's
 * hb*hb - a*c form cancels catastrophically in fp32 for the r=1000 ground spheres. */
typedef struct { float t; int prim; } hit32_t;

static inline int tri_test32(const prim32_t* p, const float* o, const float* d, float tmin, float tmax, float* tout) {
    float hx = d[1] * p->e2[2] - d[2] * p->e2[1];
    float hy = d[2] * p->e2[0] - d[0] * p->e2[2];
    float hz = d[0] * p->e2[1] - d[1] * p->e2[0];
    float det = p->e1[0] * hx + p->e1[1] * hy + p->e1[2] * hz;
    if (fabsf(det) < 1e-10f) return 0;
    float f = 1.0f / det;
    float sx = o[0] - p->a[0], sy = o[1] - p->a[1], sz = o[2] - p->a[2];
    float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < 0.0f || u > 1.0f) return 0;
    float qx = sy * p->e1[2] - sz * p->e1[1];
    float qy = sz * p->e1[0] - sx * p->e1[2];
    float qz = sx * p->e1[1] - sy * p->e1[0];
    float v = f * (d[0] * qx + d[1] * qy + d[2] * qz);
    if (v < 0.0f || u + v > 1.0f) return 0;
    float t = f * (p->e2[0] * qx + p->e2[1] * qy + p->e2[2] * qz);
    if (t < tmin || t > tmax) return 0;
    *tout = t;
    return 1;
}
static inline int sphere_test32(const prim32_t* p, const float* o, const float* d, float tmin, float tmax, float* tout) {
    float a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (!(a > 1e-18f)) return 0;
    float inva = 1.0f / a;
    float ocx = o[0] - p->c[0], ocy = o[1] - p->c[1], ocz = o[2] - p->c[2];
    float hb = ocx * d[0] + ocy * d[1] + ocz * d[2];
    float k = hb * inva;
    float lx = ocx - d[0] * k, ly = ocy - d[1] * k, lz = ocz - d[2] * k;
    float disc = a * (p->r2 - (lx * lx + ly * ly + lz * lz));
    if (disc < 0.0f) return 0;
    float sq = sqrtf(disc);
    float root = (-hb - sq) * inva;
    if (root < tmin || root > tmax) {
        root = (-hb + sq) * inva;
        if (root < tmin || root > tmax) return 0;
    }
    *tout = root;
    return 1;
}
static inline int prim_test32(const prim32_t* p, const float* o, const float* d, float tmin, float tmax, float* t) {
    return p->kind == ORC_TRIANGLE ? tri_test32(p, o, d, tmin, tmax, t) : sphere_test32(p, o, d, tmin, tmax, t);
}

typedef struct { float inv[3]; int z[3]; } slab32;
static inline slab32 slab_prep(const float* d) {
    slab32 s;
    for (int k = 0; k < 3; k++) {
        s.z[k] = fabsf(d[k]) < 1e-15f;
        s.inv[k] = s.z[k] ? 0.0f : 1.0f / d[k];
    }
    return s;
}
/* HitAccel in fp32 over outward-rounded boxes; inclusive at equality so that an
 * equal-t candidate in a later box is still examined by the tie rule. */
static inline int box_test32(const node32_t* n, const float* o, const slab32* sl, float tmin, float tmax) {
    float near = tmin;
    for (int k = 0; k < 3; k++) {
        if (sl->z[k]) {
            if (o[k] < n->mn[k] || o[k] > n->mx[k]) return 0;
        } else {
            float t0 = (n->mn[k] - o[k]) * sl->inv[k], t1 = (n->mx[k] - o[k]) * sl->inv[k];
            if (sl->inv[k] < 0.0f) { float t = t0; t0 = t1; t1 = t; }
            if (t0 > near) near = t0;
            if (t1 < tmax) tmax = t1;
            if (tmax < near) return 0;
        }
    }
    return 1;
}

/* Threaded stackless closest hit on the reference layout (gpu_opencl_windows.go:758-777). */
static hit32_t closest32(const orc_scene* s, const float* o, const float* d, float tmin, float tmax,
                         uint32_t* nv, uint32_t* pt) {
    slab32 sl = slab_prep(d);
    hit32_t best = {tmax, -1};
    uint32_t visits = 0, tests = 0;
    int ni = 0;
    while (ni >= 0 && ni < s->nnode) {
        const node32_t* n = &s->n32[ni];
        visits++;
        if (!box_test32(n, o, &sl, tmin, best.t)) { ni = n->escape; continue; }
        uint32_t meta = n->meta;
        if (meta & 0x80000000u) {
            int count = (int)(((meta >> 27) & 15u) + 1u), start = (int)(meta & 0x07ffffffu);
            for (int i = 0; i < count; i++) {
                int pid = s->order[start + i];
                float t;
                tests++;
                if (prim_test32(&s->p32[pid], o, d, tmin, best.t, &t)) {
                    if (best.prim < 0 || t < best.t || pid < best.prim) { best.t = t; best.prim = pid; }
                }
            }
            ni = n->escape;
        } else {
            ni = (int)(meta & 0x07ffffffu);
        }
    }
    if (nv) *nv = visits;
    if (pt) *pt = tests;
    return best;
}
static int any32(const orc_scene* s, const float* o, const float* d, float tmin, float tmax, uint32_t* nv, uint32_t* pt) {
    slab32 sl = slab_prep(d);
    uint32_t visits = 0, tests = 0;
    int ni = 0, hit = 0;
    while (ni >= 0 && ni < s->nnode) {
        const node32_t* n = &s->n32[ni];
        visits++;
        if (!box_test32(n, o, &sl, tmin, tmax)) { ni = n->escape; continue; }
        uint32_t meta = n->meta;
        if (meta & 0x80000000u) {
            int count = (int)(((meta >> 27) & 15u) + 1u), start = (int)(meta & 0x07ffffffu);
            for (int i = 0; i < count && !hit; i++) {
                float t;
                tests++;
                hit = prim_test32(&s->p32[s->order[start + i]], o, d, tmin, tmax, &t);
            }
            if (hit) break;
            ni = n->escape;
        } else {
            ni = (int)(meta & 0x07ffffffu);
        }
    }
    if (nv) *nv = visits;
    if (pt) *pt = tests;
    return hit;
}

/* ------------------------------------------------------------------ thread pool helper */
typedef struct {
    int64_t n, chunk;
    _Atomic int64_t next;
    void (*fn)(void* ctx, int64_t b, int64_t e);
    void* ctx;
} par_t;
static void* par_worker(void* arg) {
    par_t* p = (par_t*)arg;
    for (;;) {
        int64_t b = atomic_fetch_add(&p->next, p->chunk);
        if (b >= p->n) break;
        int64_t e = b + p->chunk < p->n ? b + p->chunk : p->n;
        p->fn(p->ctx, b, e);
    }
    return NULL;
}
static int ncpu(int req) {
    if (req > 0) return req;
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    return c > 0 ? (int)c : 1;
}
static void parallel_for(int64_t n, int64_t chunk, int nthreads, void (*fn)(void*, int64_t, int64_t), void* ctx) {
    par_t p;
    p.n = n; p.chunk = chunk > 0 ? chunk : 1; p.fn = fn; p.ctx = ctx;
    atomic_init(&p.next, 0);
    int nt = ncpu(nthreads);
    if (nt > 256) nt = 256;
    if (nt <= 1 || n <= p.chunk) { par_worker(&p); return; }
    pthread_t th[256];
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, par_worker, &p);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
}

typedef struct {
    const orc_scene* s; const float* rays; int32_t* prim; float* t; uint32_t* nv; uint32_t* pt; int any;
} tr32_ctx;
static void tr32_fn(void* c, int64_t b, int64_t e) {
    tr32_ctx* x = (tr32_ctx*)c;
    for (int64_t i = b; i < e; i++) {
        const float* r = x->rays + 8 * i;
        uint32_t nv = 0, pt = 0;
        if (x->any) {
            x->prim[i] = any32(x->s, r, r + 4, r[3], r[7], &nv, &pt);
        } else {
            hit32_t h = closest32(x->s, r, r + 4, r[3], r[7], &nv, &pt);
            x->prim[i] = h.prim;
            x->t[i] = h.prim >= 0 ? h.t : 0.0f;
        }
        if (x->nv) x->nv[i] = nv;
        if (x->pt) x->pt[i] = pt;
    }
}
void orc_hit32(const orc_scene* s, int64_t n, const float* rays, int32_t* prim, float* t,
               uint32_t* nv, uint32_t* pt, int nthreads) {
    tr32_ctx c = {s, rays, prim, t, nv, pt, 0};
    parallel_for(n, 4096, nthreads, tr32_fn, &c);
}
void orc_any32(const orc_scene* s, int64_t n, const float* rays, int32_t* occ, uint32_t* nv, uint32_t* pt, int nthreads) {
    tr32_ctx c = {s, rays, occ, NULL, nv, pt, 1};
    parallel_for(n, 4096, nthreads, tr32_fn, &c);
}
void orc_brute32(const orc_scene* s, int64_t n, const float* rays, int32_t* prim, float* t) {
    for (int64_t k = 0; k < n; k++) {
        const float* r = rays + 8 * k;
        float best = r[7];
        int bp = -1;
        for (int i = 0; i < s->nprim; i++) {
            float tt;
            if (prim_test32(&s->p32[i], r, r + 4, r[3], best, &tt)) {
                if (bp < 0 || tt < best || i < bp) { best = tt; bp = i; }
            }
        }
        prim[k] = bp;
        t[k] = bp >= 0 ? best : 0.0f;
    }
}

typedef struct { const orc_scene* s; const float* rays; int any; int32_t* out; } t64_ctx;
static void t64_fn(void* c, int64_t b, int64_t e) {
    t64_ctx* x = (t64_ctx*)c;
    for (int64_t i = b; i < e; i++) {
        const float* r = x->rays + 8 * i;
        v3 o = V(r[0], r[1], r[2]), d = V(r[4], r[5], r[6]);
        int res;
        if (x->any) {
            res = scene_occluded64(x->s, o, d, r[3], r[7]);
        } else {
            hit64 h;
            res = scene_hit64(x->s, o, d, r[3], r[7], &h) ? h.prim : -1;
        }
        if (x->out) x->out[i] = res;
    }
}
double orc_time_hit64(const orc_scene* s, int64_t n, const float* rays, int any_hit, int nthreads, int32_t* out) {
    t64_ctx c = {s, rays, any_hit, out};
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    parallel_for(n, 1024, nthreads, t64_fn, &c);
    clock_gettime(CLOCK_MONOTONIC, &b);
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

/* ------------------------------------------------------------------ sampling */
/* radicalInverse, renderer.go:1844-1858 */
double orc_radical_inverse(uint64_t n, uint64_t base) {
    if (base == 2) {
        uint64_t r = n;
        r = ((r >> 1) & 0x5555555555555555ull) | ((r & 0x5555555555555555ull) << 1);
        r = ((r >> 2) & 0x3333333333333333ull) | ((r & 0x3333333333333333ull) << 2);
        r = ((r >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((r & 0x0F0F0F0F0F0F0F0Full) << 4);
        r = ((r >> 8) & 0x00FF00FF00FF00FFull) | ((r & 0x00FF00FF00FF00FFull) << 8);
        r = ((r >> 16) & 0x0000FFFF0000FFFFull) | ((r & 0x0000FFFF0000FFFFull) << 16);
        r = (r >> 32) | (r << 32);
        return (double)r * 5.421010862427522e-20;
    }
    double invBase = 1.0 / (double)base, inv = invBase, result = 0.0;
    while (n > 0) {
        uint64_t digit = n % base;
        result += (double)digit * inv;
        n /= base;
        inv *= invBase;
    }
    return result;
}
/* pixelSampleRotation, renderer.go:1906-1915 */
void orc_pixel_rotation(int x, int y, double* rx, double* ry) {
    uint32_t h1 = (uint32_t)x * 0x9e3779b1u ^ (uint32_t)y * 0x85ebca77u;
    h1 ^= h1 >> 16;
    h1 *= 0x7feb352du;
    h1 ^= h1 >> 15;
    uint32_t h2 = h1 ^ 0x68bc21ebu;
    h2 ^= h2 >> 13;
    const double inv = 1.0 / 4294967296.0;
    *rx = (double)h1 * inv;
    *ry = (double)h2 * inv;
}
/* powerHeuristic, renderer.go:1395-1402 */
double orc_power_heuristic(double a, double b) {
    double a2 = a * a, b2 = b * b;
    if (a2 + b2 <= 1e-20) return 0;
    return a2 / (a2 + b2);
}
/* Counter-based generator replacing math/rand (renderer.go:2741).  The mixer is the
 * "lowbias32" integer hash (C. Wellons, hash-prospector, public domain). */
static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
uint32_t orc_path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    return mix32(seed ^ mix32(pixel ^ mix32(sample + 0x9e3779b9u)));
}
uint32_t orc_rng_u32(uint32_t key, uint32_t dim) { return mix32(key ^ mix32(dim * 0x85ebca6bu + 0x632be5abu)); }
float orc_rng_float(uint32_t key, uint32_t dim) { return (float)(orc_rng_u32(key, dim) >> 8) * (1.0f / 16777216.0f); }

/* Camera.basis, renderer.go:947-960 */
void orc_camera_basis(const double* pos, const double* target, const double* up, double fov, double aspect, double* o12) {
    double theta = fov * M_PI / 180;
    double h = tan(theta / 2);
    double vh = 2 * h, vw = aspect * vh;
    v3 P = V(pos[0], pos[1], pos[2]), T = V(target[0], target[1], target[2]), U = V(up[0], up[1], up[2]);
    v3 w = unit(sub(P, T));
    v3 u = unit(cross(U, w));
    v3 v = cross(w, u);
    v3 horiz = mul(u, vw), vert = mul(v, vh);
    v3 lower = sub(sub(sub(P, mul(horiz, .5)), mul(vert, .5)), w);
    double* o = o12;
    o[0] = P.x; o[1] = P.y; o[2] = P.z;
    o[3] = lower.x; o[4] = lower.y; o[5] = lower.z;
    o[6] = horiz.x; o[7] = horiz.y; o[8] = horiz.z;
    o[9] = vert.x; o[10] = vert.y; o[11] = vert.z;
}

/* ------------------------------------------------------------------ integrator
 * The same source is instantiated for double and float (oracle_trace.inc). */
#define DIM_CAM_X 0u
#define DIM_CAM_Y 1u
#define DIM_BOUNCE0 16u
#define DIMS_PER_BOUNCE 64u
#define D_LIGHT 0u
#define D_GLASS 1u
#define D_RR 2u
#define D_TRI 3u
#define D_SPH 5u
#define D_COS 21u
#define D_METAL 37u
#define NTRY_SPH 8
#define NTRY_COS 8
#define NTRY_METAL 9
/* resampled direct lighting at a vertex of bounce b: the vertex key is orc_rng_u32(path key,
 * DIM_RIS0 + b); candidate k draws dims k * RIS_DIMS + o of the vertex key, o = 0 component
 * (fewer candidates than components only), 1 light pick, 2-3 triangle point (2: the sky sample's
 * squared disk radius), 4-19 sphere point / sky angle tries, 20 reservoir update */
#define DIM_RIS0 0x100000u
#define RIS_DIMS 32u
#define RIS_MAX 64
/* spatial reuse: neighbour i draws dims DIM_SPATIAL0 + 4 i + o, o = 0 x offset, 1 y offset
 * (integer hashes), 2 reservoir update (i = 0: the pixel's own reservoir) */
#define DIM_SPATIAL0 0x200000u

/* sun direction of a render description, normalised in double precision and rounded to float */
static inline void orc_sun_dir(const orc_render_desc* d, float out[3]) {
    double x = d->sun_dir[0], y = d->sun_dir[1], z = d->sun_dir[2];
    double l = sqrt(x * x + y * y + z * z);
    if (!(l > 0)) { out[0] = out[1] = out[2] = 0.0f; return; }
    out[0] = (float)(x / l); out[1] = (float)(y / l); out[2] = (float)(z / l);
}

static inline double rnd64(uint32_t key, uint32_t dim) { return (double)(orc_rng_u32(key, dim) >> 8) * (1.0 / 16777216.0); }
static inline float rnd32(uint32_t key, uint32_t dim) { return (float)(orc_rng_u32(key, dim) >> 8) * (1.0f / 16777216.0f); }

/* Camera jitter of sample smp at pixel (x0, y) of a W-wide image (renderer.go:1844-1915,
 * 1943-1950): Halton (bases 2, 3) or R2 with the pixel's Cranley-Patterson rotation, or two
 * counter-based uniforms for the random sampler; both in [0, 1). */
void orc_camera_jitter(int sampler, uint32_t seed, int W, int x0, int y, int smp, double* jx, double* jy) {
    double rotX = 0, rotY = 0;
    orc_pixel_rotation(x0, y, &rotX, &rotY);
    if (sampler == 1) {
        double hx = orc_radical_inverse((uint64_t)smp + 1, 2), hy = orc_radical_inverse((uint64_t)smp + 1, 3);
        *jx = hx + rotX; *jx = *jx - floor(*jx);
        *jy = hy + rotY; *jy = *jy - floor(*jy);
    } else if (sampler == 2) {
        const double g = 1.324717957244746, a1 = 1.0 / g, a2 = 1.0 / (g * g);
        double sx = 0.5 + (double)(smp + 1) * a1, sy = 0.5 + (double)(smp + 1) * a2;
        sx = sx - floor(sx); sy = sy - floor(sy);
        *jx = sx + rotX; *jx = *jx - floor(*jx);
        *jy = sy + rotY; *jy = *jy - floor(*jy);
    } else {
        uint32_t key = orc_path_key(seed, (uint32_t)y * (uint32_t)W + (uint32_t)x0, (uint32_t)smp);
        *jx = rnd64(key, DIM_CAM_X);
        *jy = rnd64(key, DIM_CAM_Y);
    }
}

void orc_cos_hemisphere64(const double* nn, uint32_t key, uint32_t dim0, double* out) {
    /* randCosHemisphere, renderer.go:1290-1311, with the integrator's bounded retries */
    v3 n = V(nn[0], nn[1], nn[2]);
    double x = 0, y = 0, s = 0;
    int ok = 0;
    for (int k = 0; k < NTRY_COS && !ok; k++) {
        x = rnd64(key, dim0 + 2u * (uint32_t)k) * 2 - 1;
        y = rnd64(key, dim0 + 2u * (uint32_t)k + 1u) * 2 - 1;
        s = x * x + y * y;
        ok = s < 1;
    }
    if (!ok) { x = 0; y = 0; s = 0; }
    double z = sqrt(fmax(0, 1 - s));
    double sign = n.z < 0 ? -1.0 : 1.0;
    double a = -1.0 / (sign + n.z);
    double b = n.x * n.y * a;
    v3 t = V(1 + sign * n.x * n.x * a, sign * b, -sign * n.x);
    v3 bt = V(b, sign + n.y * n.y * a, -n.y);
    v3 r = add(add(mul(t, x), mul(bt, y)), mul(n, z));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* fp64 instantiation: Go semantics */
#define R double
#define SUF(x) x##_64
#define RND rnd64
#define SQRT sqrt
#define FABS fabs
#define IS64 1
#include "oracle_trace.inc"
#undef R
#undef SUF
#undef RND
#undef SQRT
#undef FABS
#undef IS64

/* fp32 instantiation: GPU arithmetic */
#define R float
#define SUF(x) x##_32
#define RND rnd32
#define SQRT sqrtf
#define FABS fabsf
#define IS64 0
#include "oracle_trace.inc"

int64_t orc_render64(const orc_scene* s, const orc_render_desc* d, int y0, int rows, double* rgba,
                     uint32_t* guides, int64_t* shadow_rays, int nthreads) {
    return orc_render_64(s, d, y0, rows, rgba, guides, shadow_rays, nthreads);
}
int64_t orc_render32(const orc_scene* s, const orc_render_desc* d, int y0, int rows, float* rgba,
                     uint32_t* guides, int64_t* shadow_rays, int nthreads) {
    return orc_render_32(s, d, y0, rows, rgba, guides, shadow_rays, nthreads);
}

/* ------------------------------------------------------------------ single-primitive entry points
 * For known-answer tests of the primitive tests and the guide packing on their own (test
 * infrastructure, like the rest of this file). */

/* triHit (renderer.go:684-711) in fp64 on A, E1, E2 (9 doubles); ray6 = o.xyz, d.xyz */
int orc_tri_hit64(const double* g9, const double* ray6, double tmin, double tmax, double* t_out) {
    prim_t p;
    memset(&p, 0, sizeof p);
    p.kind = ORC_TRIANGLE;
    p.a = V(g9[0], g9[1], g9[2]);
    p.e1 = V(g9[3], g9[4], g9[5]);
    p.e2 = V(g9[6], g9[7], g9[8]);
    p.n = unit(cross(p.e1, p.e2));
    hit64 h;
    int r = tri_hit(&p, V(ray6[0], ray6[1], ray6[2]), V(ray6[3], ray6[4], ray6[5]), tmin, tmax, &h);
    if (t_out) *t_out = r ? h.t : 0;
    return r;
}
/* sphereHitAccel (renderer.go:650-678) in fp64; c4 = center.xyz, radius */
int orc_sphere_hit64(const double* c4, const double* ray6, double tmin, double tmax, double* t_out) {
    prim_t p;
    memset(&p, 0, sizeof p);
    p.kind = ORC_SPHERE;
    p.center = V(c4[0], c4[1], c4[2]);
    p.radius = c4[3];
    rayacc ar = prepare(V(ray6[0], ray6[1], ray6[2]), V(ray6[3], ray6[4], ray6[5]));
    hit64 h;
    int r = sphere_hit(&p, &ar, tmin, tmax, &h);
    if (t_out) *t_out = r ? h.t : 0;
    return r;
}
/* the fp32 triangle test of the GPU traversal on A, E1, E2 (9 floats) */
int orc_tri_test32(const float* g9, const float* o, const float* d, float tmin, float tmax, float* t_out) {
    prim32_t p;
    memset(&p, 0, sizeof p);
    p.kind = ORC_TRIANGLE;
    for (int k = 0; k < 3; k++) { p.a[k] = g9[k]; p.e1[k] = g9[3 + k]; p.e2[k] = g9[6 + k]; }
    float t = 0;
    int r = tri_test32(&p, o, d, tmin, tmax, &t);
    if (t_out) *t_out = r ? t : 0.0f;
    return r;
}
/* the fp32 sphere test of the GPU traversal; c4 = center.xyz, radius (r^2 rounded from fp64 as
 * the record packer does) */
int orc_sphere_test32(const double* c4, const float* o, const float* d, float tmin, float tmax, float* t_out) {
    prim32_t p;
    memset(&p, 0, sizeof p);
    p.kind = ORC_SPHERE;
    for (int k = 0; k < 3; k++) p.c[k] = (float)c4[k];
    p.r = (float)c4[3];
    p.r2 = (float)(c4[3] * c4[3]);
    p.invr = (float)(1.0 / c4[3]);
    float t = 0;
    int r = sphere_test32(&p, o, d, tmin, tmax, &t);
    if (t_out) *t_out = r ? t : 0.0f;
    return r;
}
/* packGuideNormalMat / packJitter (gpu_opencl_windows.go:787-797) as the integrator writes them */
uint32_t orc_pack_guide32(const float* n3, int mat) {
    vec_32 n = {n3[0], n3[1], n3[2]};
    return pack_guide_32(n, mat);
}
uint32_t orc_pack_jitter(float ju, float jv) { return pack_jitter_32(ju, jv); }
