/* Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the host
 * half of the C ABI (octree, keys, masks, immersed-boundary weights, kernel
 * tables) and the FP64 oracle, linked into one executable by
 * tests/test_sanitizers.py.  Exercises tree construction with refinement,
 * correction blocks, masks and export, the output and restart files, the
 * immersed-boundary stencil, and
 * the oracle's Laplace and integrating-factor solves and IB transfers on the
 * exported trees.  Exit status 0 when every call succeeds; the sanitizers
 * abort on the first error they find. */
#include "iblgf_gpu.h"
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CHECK(c)                                                   \
    do                                                             \
    {                                                              \
        if (!(c))                                                  \
        {                                                          \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            exit(1);                                               \
        }                                                          \
    } while (0)

typedef struct
{
    int       nb;
    uint64_t* key;
    int *     level, *coord, *bbase, *parent, *children, *neighbor, *influence, *influence_num;
    uint32_t* flags;
    uint64_t *mask_src, *mask_tgt;
} table;

static void export_table(const iblgf_tree* t, table* a, oracle_tree* o)
{
    int    nb, base, depth, ext, nl;
    double dx;
    CHECK(iblgf_tree_info(t, &nb, &base, &depth, &ext, &dx, &nl) == IBLGF_OK);
    a->nb = nb;
    a->key = malloc(sizeof(uint64_t) * nb);
    a->level = malloc(sizeof(int) * nb);
    a->coord = malloc(sizeof(int) * 3 * nb);
    a->bbase = malloc(sizeof(int) * 3 * nb);
    a->flags = malloc(sizeof(uint32_t) * nb);
    a->parent = malloc(sizeof(int) * nb);
    a->children = malloc(sizeof(int) * 8 * nb);
    a->neighbor = malloc(sizeof(int) * 27 * nb);
    a->influence = malloc(sizeof(int) * 189 * nb);
    a->influence_num = malloc(sizeof(int) * nb);
    a->mask_src = malloc(sizeof(uint64_t) * nb);
    a->mask_tgt = malloc(sizeof(uint64_t) * nb);
    CHECK(iblgf_tree_export(t, a->key, a->level, a->coord, a->bbase, a->flags, a->parent, a->children, a->neighbor,
              a->influence, a->influence_num, a->mask_src, a->mask_tgt) == IBLGF_OK);
    o->nb = nb;
    o->n = ext + 2;
    o->base_level = base;
    o->depth = depth;
    o->dx_base = dx;
    o->level = a->level;
    o->coord = a->coord;
    o->bbase = a->bbase;
    o->flags = a->flags;
    o->parent = a->parent;
    o->children = a->children;
    o->neighbor = a->neighbor;
    o->influence = a->influence;
    o->influence_num = a->influence_num;
    o->mask_src = a->mask_src;
    o->mask_tgt = a->mask_tgt;
}

static void free_table(table* a)
{
    free(a->key);
    free(a->level);
    free(a->coord);
    free(a->bbase);
    free(a->flags);
    free(a->parent);
    free(a->children);
    free(a->neighbor);
    free(a->influence);
    free(a->influence_num);
    free(a->mask_src);
    free(a->mask_tgt);
}

/* refine blocks that touch a ball of radius 10 base cells about the origin */
static int near_origin(void* user, const int b[3], int ext, int rl, int diff)
{
    const double s = ldexp(1.0, -rl);
    double       d2 = 0.0;
    int          k;
    (void)user;
    (void)diff;
    for (k = 0; k < 3; ++k)
    {
        const double lo = b[k] * s, hi = (b[k] + ext) * s;
        const double c = lo > 0 ? lo : (hi < 0 ? hi : 0.0);
        d2 += c * c;
    }
    return d2 < 100.0;
}

static double* random_source(const oracle_tree* o, unsigned seed)
{
    const size_t n3 = (size_t)o->n * o->n * o->n, total = n3 * o->nb;
    double*      s = malloc(sizeof(double) * total);
    size_t       i;
    srand(seed);
    for (i = 0; i < total; ++i) s[i] = (double)rand() / RAND_MAX - 0.5;
    return s;
}

static void solve_both(const oracle_tree* o, const double* g, int nmax)
{
    const size_t n3 = (size_t)o->n * o->n * o->n, total = n3 * o->nb;
    double*      src = random_source(o, 7);
    double*      dst = calloc(total, sizeof(double));
    size_t       i;
    CHECK(oracle_apply_lgf(o, g, nmax, src, dst, 0, 5) == 0);
    for (i = 0; i < total; ++i) CHECK(isfinite(dst[i]));
    memset(dst, 0, sizeof(double) * total);
    CHECK(oracle_apply_if(o, src, dst, 0.4, 0, 5) == 0);
    for (i = 0; i < total; ++i) CHECK(isfinite(dst[i]));
    free(src);
    free(dst);
}

int main(void)
{
    const int nmax = 100;
    double*   g = malloc(sizeof(double) * (nmax + 1) * (nmax + 2) * (nmax + 3) / 6);
    oracle_lgf_table(nmax, g);

    /* kernel tables and keys */
    {
        double   tab[(12 + 1) * (12 + 2) * (12 + 3) / 6], A[10 * 20];
        uint64_t infl[216];
        int      xyz[3];
        CHECK(iblgf_lgf_table(12, tab) == IBLGF_OK);
        CHECK(fabs(tab[0] - g[0]) < 1e-12);
        CHECK(iblgf_lagrange_matrix(10, 10, 0, A) == IBLGF_OK);
        CHECK(iblgf_lagrange_matrix(10, 5, 1, A) == IBLGF_OK);
        const uint64_t k = iblgf_key_make(3, 5, 6, 4);
        CHECK(iblgf_key_coord(k, xyz) == 4 && xyz[0] == 3 && xyz[1] == 5 && xyz[2] == 6);
        CHECK(iblgf_key_parent(iblgf_key_child(k, 5)) == k);
        CHECK(iblgf_key_influence(k, infl) <= 189);
        CHECK(isfinite(iblgf_lgf_get(120, 3, 1)) && isfinite(iblgf_if_kernel(2, 1, 0, 0.7)));
    }

    /* uniform single-level tree (4^3 blocks of 8^3) */
    {
        const int   base[3] = {-16, -16, -16}, ext[3] = {32, 32, 32}, region[6] = {-16, -16, -16, 32, 32, 32};
        iblgf_tree* t = NULL;
        table       a;
        oracle_tree o;
        CHECK(iblgf_tree_create(base, ext, 8, 1.0 / 32, 1, region, &t) == IBLGF_OK);
        CHECK(iblgf_tree_init_refine(t, 0, 0, NULL, NULL) == IBLGF_OK);
        CHECK(iblgf_tree_build_masks(t, 0) == IBLGF_OK);
        export_table(t, &a, &o);
        solve_both(&o, g, nmax);
        free_table(&a);
        iblgf_tree_destroy(t);
    }

    /* refined tree: two levels around the origin, correction blocks, masks */
    {
        const int   base[3] = {-32, -32, -32}, ext[3] = {64, 64, 64}, region[6] = {-16, -16, -16, 32, 32, 32};
        iblgf_tree* t = NULL;
        table       a;
        oracle_tree o;
        CHECK(iblgf_tree_create(base, ext, 8, 1.0 / 32, 1, region, &t) == IBLGF_OK);
        CHECK(iblgf_tree_init_refine(t, 2, 0, near_origin, NULL) == IBLGF_OK);
        CHECK(iblgf_tree_build_masks(t, 0) == IBLGF_OK);
        export_table(t, &a, &o);
        CHECK(o.depth > o.base_level + 1);
        solve_both(&o, g, nmax);
        /* output and restart files: Chombo HDF5 and Tree::write, read back */
        {
            const size_t  nv = (size_t)o.nb * o.n * o.n * o.n;
            double*       u = (double*)malloc(3 * nv * sizeof(double));
            double*       back = (double*)calloc(3 * nv, sizeof(double));
            const char*   names[1] = {"u"};
            const int     nc[1] = {3};
            const double* data[1] = {u};
            iblgf_tree*   t2 = NULL;
            int           nb2 = 0, b2, d2, e2, nl2;
            double        dx2;
            for (size_t i = 0; i < 3 * nv; ++i) u[i] = (double)(i % 97) - 48.0;
            CHECK(iblgf_chombo_write(t, "sanitize_flow.hdf5", 1, names, nc, data, 1, 1, 0.0, 1.0) == IBLGF_OK);
            CHECK(iblgf_chombo_read(t, "sanitize_flow.hdf5", "u", 3, IBLGF_MESH_FACE, back) == IBLGF_OK);
            CHECK(iblgf_chombo_read(t, "sanitize_flow.hdf5", "w", 3, IBLGF_MESH_FACE, back) == IBLGF_ERR_ARG);
            CHECK(iblgf_tree_write_restart(t, "sanitize_tree.bin") == IBLGF_OK);
            CHECK(iblgf_tree_read_restart("sanitize_tree.bin", base, ext, 8, 1.0 / 32, &t2) == IBLGF_OK);
            CHECK(iblgf_tree_info(t2, &nb2, &b2, &d2, &e2, &dx2, &nl2) == IBLGF_OK && nb2 == o.nb);
            iblgf_tree_destroy(t2);
            remove("sanitize_flow.hdf5");
            remove("sanitize_tree.bin");
            free(u);
            free(back);
        }
        free_table(&a);
        iblgf_tree_destroy(t);
    }

    /* immersed boundary: 200 points on a sphere, roma delta */
    {
        const int   base[3] = {-32, -32, -32}, ext[3] = {64, 64, 64}, region[6] = {-16, -16, -16, 32, 32, 32};
        const int   np = 200;
        double*     pts = malloc(sizeof(double) * 3 * np);
        iblgf_tree* t = NULL;
        table       a;
        oracle_tree o;
        oracle_ib   ib;
        long        sz[4];
        int         i;
        for (i = 0; i < np; ++i)
        {
            const double z = 1.0 - 2.0 * (i + 0.5) / np, r = sqrt(1.0 - z * z), ph = 2.399963229728653 * i;
            pts[3 * i] = 0.15 * r * cos(ph);
            pts[3 * i + 1] = 0.15 * r * sin(ph);
            pts[3 * i + 2] = 0.15 * z;
        }
        CHECK(iblgf_tree_create(base, ext, 8, 1.0 / 32, 1, region, &t) == IBLGF_OK);
        CHECK(iblgf_tree_set_ib(t, pts, np, 1, IBLGF_DELTA_ROMA, 4.0) == IBLGF_OK);
        CHECK(iblgf_tree_init_refine(t, 1, 0, NULL, NULL) == IBLGF_OK);
        CHECK(iblgf_tree_build_masks(t, 0) == IBLGF_OK);
        CHECK(iblgf_tree_ib_stencil(t, sz, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL) == IBLGF_OK);
        {
            int*    ip = malloc(sizeof(int) * (3 * np + 1));
            int*    in = malloc(sizeof(int) * (sz[0] + 1));
            double* iw = malloc(sizeof(double) * (sz[0] + 1));
            int*    sp = malloc(sizeof(int) * (sz[1] + 1));
            int*    sd = malloc(sizeof(int) * 2 * (sz[1] + 1));
            int*    sc = malloc(sizeof(int) * (sz[2] + 1));
            double* sw = malloc(sizeof(double) * (sz[2] + 1));
            int*    ow = malloc(sizeof(int) * np);
            CHECK(iblgf_tree_ib_stencil(t, sz, ip, in, iw, sp, sd, sc, sw, ow) == IBLGF_OK);
            free(ip);
            free(in);
            free(iw);
            free(sp);
            free(sd);
            free(sc);
            free(sw);
            free(ow);
        }
        export_table(t, &a, &o);
        ib.npoints = np;
        ib.x = pts;
        ib.ib_level = 1;
        ib.delta = 2;
        ib.ddf_radius = 4.5;
        ib.safety_dis = 4.0;
        {
            const size_t n3 = (size_t)o.n * o.n * o.n, total = n3 * o.nb;
            double*      f0 = calloc(total, sizeof(double));
            double*      f1 = calloc(total, sizeof(double));
            double*      f2 = calloc(total, sizeof(double));
            double*      force = malloc(sizeof(double) * 3 * np);
            double*      back = calloc(3 * np, sizeof(double));
            int*         owned = malloc(sizeof(int) * np);
            for (i = 0; i < 3 * np; ++i) force[i] = (i % 7) * 0.1 - 0.3;
            oracle_ib_smear(&o, &ib, force, f0, f1, f2);
            oracle_ib_project(&o, &ib, f0, f1, f2, back);
            oracle_ib_owned(&o, &ib, owned);
            for (i = 0; i < 3 * np; ++i) CHECK(isfinite(back[i]));
            free(f0);
            free(f1);
            free(f2);
            free(force);
            free(back);
            free(owned);
        }
        free_table(&a);
        iblgf_tree_destroy(t);
        free(pts);
    }

    /* argument errors on the host entry points */
    CHECK(iblgf_tree_build_masks(NULL, 0) == IBLGF_ERR_ARG);
    CHECK(iblgf_lagrange_matrix(18, 5, 3, NULL) == IBLGF_ERR_ARG);
    CHECK(strlen(iblgf_last_error()) > 0);
    free(g);
    printf("sanitized host checks passed\n");
    return 0;
}
