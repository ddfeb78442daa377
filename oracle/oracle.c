/* CPU FP64 oracle of the LGF fast-multipole path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Restates, in plain C and in the reference's own loop order:
 *   LGF table / asymptotics  include/iblgf/lgf/lgf_gl_lookup.hpp:33-52, 74-119
 *                            include/iblgf/lgf/lgf_table_gen/lgf_integrator.py:7-25
 *   IF kernel                include/iblgf/lgf/lgf_ge.hpp:83, 158-168
 *   spectrum cache           include/iblgf/lgf/lgf.hpp:77-98, 221-262
 *   block convolution        include/iblgf/utilities/convolution.hpp:226-304,
 *                            src/utilities/convolution.cpp:357-570 (pruned legs)
 *   Nli                      include/iblgf/fmm/fmm_nli.hpp:101-247, 319-374
 *   Fmm::apply               include/iblgf/fmm/fmm.hpp:729-870, 988-1195
 *   cell_center_nli          include/iblgf/interpolation/cell_center_nli_intrp.hpp:171-238, 310-420, 557-787
 *   halo exchange            include/iblgf/domain/mpi/haloCommunicator.hpp:67-128
 *   PoissonSolver            include/iblgf/solver/poisson/poisson.hpp:427-494, 587-894, 896-1194
 *   2-D Helmholtz LGF        include/iblgf/lgf/helmholtz.hpp:143-153,
 *                            include/iblgf/lgf/helmholtz_lookup.hpp:60-68, 146-181
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <math.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

/* Workers of the rank-split loops: the OpenMP team, or one worker when the
 * library is built without OpenMP (every range then runs on it in turn). */
static int worker_id(void)
{
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}
static int worker_count(void)
{
#ifdef _OPENMP
    return omp_get_num_threads();
#else
    return 1;
#endif
}
static double wall_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

#define O_PI 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------------ */
/* Scaled modified Bessel functions                                          */
/* ------------------------------------------------------------------------ */
void oracle_ive(double x, int kmax, double* out)
{
    int k;
    if (x <= 0.0)
    {
        out[0] = 1.0;
        for (k = 1; k <= kmax; ++k) out[k] = 0.0;
        return;
    }
    if (x > 1.0e4)
    {
        /* Hankel expansion e^{-x} I_k(x) ~ (2 pi x)^{-1/2} sum_j (-1)^j a_j(k) / x^j */
        double pre = 1.0 / sqrt(2.0 * O_PI * x);
        for (k = 0; k <= kmax; ++k)
        {
            double mu = 4.0 * (double)k * (double)k, term = 1.0, s = 1.0;
            int    j;
            for (j = 1; j < 100; ++j)
            {
                double a = 2.0 * j - 1.0;
                term *= -(mu - a * a) / (8.0 * x * (double)j);
                s += term;
                if (fabs(term) < 1e-18 * fabs(s)) break;
            }
            out[k] = pre * s;
        }
        return;
    }
    {
        /* backward recurrence from N well past max(k, x) */
        int    N = kmax + (int)x + (int)(15.0 * sqrt(x)) + 50;
        double ip = 0.0, ic = 1e-250, norm = 0.0;
        for (k = 0; k <= kmax; ++k) out[k] = 0.0;
        norm = 2.0 * ic;
        for (k = N; k > 0; --k)
        {
            double im = ip + (2.0 * (double)k / x) * ic;
            ip = ic;
            ic = im; /* ic = I_{k-1} */
            if (k - 1 <= kmax) out[k - 1] = ic;
            norm += (k - 1 == 0) ? ic : 2.0 * ic;
            if (ic > 1e200)
            {
                int j;
                ic *= 1e-200;
                ip *= 1e-200;
                norm *= 1e-200;
                for (j = (k - 1 < 0 ? 0 : k - 1); j <= kmax; ++j) out[j] *= 1e-200;
            }
        }
        for (k = 0; k <= kmax; ++k) out[k] /= norm;
    }
}

/* ------------------------------------------------------------------------ */
/* Laplace LGF                                                                */
/* ------------------------------------------------------------------------ */
static long lgf_idx(int x, int y, int z)
{
    return (long)x * (x + 1) * (x + 2) / 6 + (long)y * (y + 1) / 2 + z;
}

void oracle_lgf_table(int nmax, double* g)
{
    /* G(n) = -int_0^inf prod_i ive(n_i, 2t) dt, on t = e^s, trapezoid with
     * step 1/32 on s in [-44, ln 1e11], end corrections and analytic tail. */
    const double h = 1.0 / 32.0, s0 = -44.0;
    const int    ns = (int)ceil((log(1e11) - s0) / h);
    const double s1 = s0 + ns * h;
    const long   size = lgf_idx(nmax, nmax, nmax) + 1;
    double*      iv = (double*)malloc(sizeof(double) * (nmax + 1));
    long         i;
    int          q, x, y, z;
    for (i = 0; i < size; ++i) g[i] = 0.0;
    for (q = 0; q <= ns; ++q)
    {
        double t = exp(s0 + q * h);
        double w = h * t * (q == ns ? 0.5 : 1.0);
        long   id = 0;
        oracle_ive(2.0 * t, nmax, iv);
        for (x = 0; x <= nmax; ++x)
            for (y = 0; y <= x; ++y)
                for (z = 0; z <= y; ++z) g[id++] += w * iv[x] * iv[y] * iv[z];
    }
    {
        double t = exp(s1), T = t;
        double ce = (h * h / 24.0 - h * h * h * h / 5760.0) * t;
        double c0 = pow(4.0 * O_PI, -1.5);
        long   id = 0;
        oracle_ive(2.0 * t, nmax, iv);
        for (x = 0; x <= nmax; ++x)
            for (y = 0; y <= x; ++y)
                for (z = 0; z <= y; ++z)
                {
                    double m = 4.0 * ((double)x * x + (double)y * y + (double)z * z) - 3.0;
                    double tail = c0 * (2.0 / sqrt(T) - m / 24.0 * pow(T, -1.5));
                    g[id] = -(g[id] + ce * iv[x] * iv[y] * iv[z] + tail);
                    ++id;
                }
    }
    free(iv);
}

double oracle_lgf_asym(int n1, int n2, int n3)
{
    double a = (double)n1 * n1, b = (double)n2 * n2, c = (double)n3 * n3;
    double r = sqrt(a + b + c);
    double r2 = r * r, r6 = r2 * r2 * r2, r7 = r6 * r, r13 = r6 * r6 * r;
    double v = -1.0 / (4.0 * O_PI * r);
    v -= (a * a + b * b + c * c - 3.0 * a * b - 3.0 * b * c - 3.0 * c * a) / (16.0 * O_PI * r7);
    if (n1 < 600 || n2 < 600 || n3 < 600)
    {
        double a3 = a * a * a, b3 = b * b * b, c3 = c * c * c;
        double p = 23.0 * (a3 * a + b3 * b + c3 * c) -
                   244.0 * (b3 * c + c3 * b + a3 * b + b3 * a + a3 * c + c3 * a) +
                   621.0 * (a * a * b * b + b * b * c * c + c * c * a * a) -
                   228.0 * (a * a * b * c + a * b * b * c + a * b * c * c);
        v += -3.0 * p / r13 / 4.0 * (8.0 / (768.0 * O_PI));
    }
    return v;
}

double oracle_lgf_get(const double* g, int nmax, int x, int y, int z)
{
    int t;
    x = abs(x);
    y = abs(y);
    z = abs(z);
    if (x <= nmax && y <= nmax && z <= nmax)
    {
        if (x < y) { t = x; x = y; y = t; }
        if (x < z) { t = x; x = z; z = t; }
        if (y < z) { t = y; y = z; z = t; }
        return g[lgf_idx(x, y, z)];
    }
    return oracle_lgf_asym(x, y, z);
}

void oracle_direct_sum(const double* g, int nmax, int ns, const int* sx, const double* q, int nt,
    const int* tx, double scale, double* phi)
{
    int i, j;
    for (i = 0; i < nt; ++i)
    {
        double acc = 0.0;
        for (j = 0; j < ns; ++j)
            acc += oracle_lgf_get(g, nmax, tx[3 * i] - sx[3 * j], tx[3 * i + 1] - sx[3 * j + 1],
                       tx[3 * i + 2] - sx[3 * j + 2]) *
                   q[j];
        phi[i] = scale * acc;
    }
}

/* ------------------------------------------------------------------------ */
/* Lagrange transfer matrices                                                 */
/* ------------------------------------------------------------------------ */
void oracle_lagrange_matrix(int n, int cap, int rp, double* A)
{
    int c, p, l;
    memset(A, 0, sizeof(double) * (size_t)n * 2 * n);
    for (c = 1; c < 2 * n - 1; ++c)
    {
        double cpos = rp ? (c - n + 0.5) : (double)(c - n + 1); /* doubled, centred position */
        int    bl, br;
        long double sum = 0.0L;
        if (!rp && c % 2 == 0)
        {
            A[(size_t)(c / 2) * 2 * n + c - 1] = 1.0;
            continue;
        }
        if ((rp ? (c - n + 0.5) : (double)(c - n)) < -1.0)
        {
            bl = (c + 1) / 2 - cap / 2;
            if (bl < 0) bl = 0;
            br = bl + cap;
            if (br > n) br = n;
        }
        else
        {
            br = (c + 1) / 2 + cap / 2;
            if (br > n) br = n;
            bl = br - cap;
            if (bl < 0) bl = 0;
        }
        for (p = bl; p < br; ++p)
        {
            long double m = 1.0L;
            int         pp = 2 * p - n + 1;
            for (l = bl; l < br; ++l)
                if (l != p) m /= -(long double)(2 * l - n + 1 - pp);
            m /= (long double)(cpos - pp);
            A[(size_t)p * 2 * n + c - 1] = (double)m;
            sum += m;
        }
        for (p = 0; p < n; ++p) A[(size_t)p * 2 * n + c - 1] /= (double)sum;
    }
    for (c = 2 * n - 1; c > n - 1; --c)
        for (p = 0; p < n; ++p) A[(size_t)p * 2 * n + c] = A[(size_t)p * 2 * n + c - 2];
}

/* n x n sub-block of the n x 2n matrix for child side b: S[o][i] = A[o][b n + i] */
static void sub_matrix(const double* A, int n, int b, double* S)
{
    int o, i;
    for (o = 0; o < n; ++o)
        for (i = 0; i < n; ++i) S[o * n + i] = A[(size_t)o * 2 * n + b * n + i];
}
static void transpose(const double* S, int n, double* T)
{
    int o, i;
    for (o = 0; o < n; ++o)
        for (i = 0; i < n; ++i) T[i * n + o] = S[o * n + i];
}

/* restriction sub-matrix (cell_center_nli simple coarsening), rp 0/1 */
static void restrict_matrix(int n, int b, int rp, double* M)
{
    int i, j;
    memset(M, 0, sizeof(double) * n * n);
    for (i = 1; i < n - 1; ++i)
    {
        /* simple matrix columns 2i-1 (and 2i for rp 1), then shift by off */
        int off = (b == 0) ? 1 : n - 1;
        for (j = 1; j < n - 1; ++j)
        {
            int col = off + j - 1;
            double v = 0.0;
            if (col == 2 * i - 1) v = rp ? 0.5 : 1.0;
            if (rp && col == 2 * i) v = 0.5;
            M[i * n + j] = v;
        }
    }
}

/* out[k][j][i] (=|+=) sum Mz[k][c] My[j][b] Mx[i][a] in[c][b][a] */
static void transfer3(int n, const double* Mx, const double* My, const double* Mz, const double* in,
    double* out, int accumulate)
{
    size_t  n3 = (size_t)n * n * n;
    double* t1 = (double*)calloc(n3, sizeof(double));
    double* t2 = (double*)calloc(n3, sizeof(double));
    int     a, b, c, i, j, k;
    for (c = 0; c < n; ++c)
        for (b = 0; b < n; ++b)
            for (i = 0; i < n; ++i)
            {
                double s = 0.0;
                for (a = 0; a < n; ++a) s += Mx[i * n + a] * in[((size_t)c * n + b) * n + a];
                t1[((size_t)c * n + b) * n + i] = s;
            }
    for (c = 0; c < n; ++c)
        for (j = 0; j < n; ++j)
            for (i = 0; i < n; ++i)
            {
                double s = 0.0;
                for (b = 0; b < n; ++b) s += My[j * n + b] * t1[((size_t)c * n + b) * n + i];
                t2[((size_t)c * n + j) * n + i] = s;
            }
    for (k = 0; k < n; ++k)
        for (j = 0; j < n; ++j)
            for (i = 0; i < n; ++i)
            {
                double s = 0.0;
                for (c = 0; c < n; ++c) s += Mz[k * n + c] * t2[((size_t)c * n + j) * n + i];
                if (accumulate) out[((size_t)k * n + j) * n + i] += s;
                else
                    out[((size_t)k * n + j) * n + i] = s;
            }
    free(t1);
    free(t2);
}

/* ------------------------------------------------------------------------ */
/* Block convolution by zero-padded DFT of size P = 2n-1                     */
/* ------------------------------------------------------------------------ */
typedef struct
{
    int     n, P, H;
    double* cw; /* cos(2 pi m / P) */
    double* sw; /* sin(2 pi m / P) */
} dft_t;

static void dft_init(dft_t* d, int n)
{
    int m;
    d->n = n;
    d->P = 2 * n - 1;
    d->H = d->P / 2 + 1;
    d->cw = (double*)malloc(sizeof(double) * d->P);
    d->sw = (double*)malloc(sizeof(double) * d->P);
    for (m = 0; m < d->P; ++m)
    {
        d->cw[m] = cos(2.0 * O_PI * m / d->P);
        d->sw[m] = sin(2.0 * O_PI * m / d->P);
    }
}
static void dft_free(dft_t* d)
{
    free(d->cw);
    free(d->sw);
}
static size_t spec_size(const dft_t* d) { return (size_t)d->H * d->P * d->P; }

/* r2c of an nin^3 block (x fastest) zero-padded to P^3.  Output (re, im) at
 * [(kz*P + ky)*H + kx].  Legs pruned to the nonzero input lines. */
static void dft_forward(const dft_t* d, const double* in, int nin, double* ore, double* oim)
{
    const int P = d->P, H = d->H;
    size_t    sa = (size_t)nin * nin * H, sb = (size_t)nin * P * H;
    double *  are = (double*)malloc(sizeof(double) * sa), *aim = (double*)malloc(sizeof(double) * sa);
    double *  bre = (double*)malloc(sizeof(double) * sb), *bim = (double*)malloc(sizeof(double) * sb);
    int       x, y, z, kx, ky, kz;
    for (z = 0; z < nin; ++z)
        for (y = 0; y < nin; ++y)
        {
            const double* line = in + ((size_t)z * nin + y) * nin;
            for (kx = 0; kx < H; ++kx)
            {
                double sr = 0.0, si = 0.0;
                for (x = 0; x < nin; ++x)
                {
                    int m = (kx * x) % P;
                    sr += line[x] * d->cw[m];
                    si -= line[x] * d->sw[m];
                }
                are[((size_t)z * nin + y) * H + kx] = sr;
                aim[((size_t)z * nin + y) * H + kx] = si;
            }
        }
    for (z = 0; z < nin; ++z)
        for (ky = 0; ky < P; ++ky)
            for (kx = 0; kx < H; ++kx)
            {
                double sr = 0.0, si = 0.0;
                for (y = 0; y < nin; ++y)
                {
                    int    m = (ky * y) % P;
                    double c = d->cw[m], s = -d->sw[m];
                    double ar = are[((size_t)z * nin + y) * H + kx], ai = aim[((size_t)z * nin + y) * H + kx];
                    sr += ar * c - ai * s;
                    si += ar * s + ai * c;
                }
                bre[((size_t)z * P + ky) * H + kx] = sr;
                bim[((size_t)z * P + ky) * H + kx] = si;
            }
    for (kz = 0; kz < P; ++kz)
        for (ky = 0; ky < P; ++ky)
            for (kx = 0; kx < H; ++kx)
            {
                double sr = 0.0, si = 0.0;
                for (z = 0; z < nin; ++z)
                {
                    int    m = (kz * z) % P;
                    double c = d->cw[m], s = -d->sw[m];
                    double br = bre[((size_t)z * P + ky) * H + kx], bi = bim[((size_t)z * P + ky) * H + kx];
                    sr += br * c - bi * s;
                    si += br * s + bi * c;
                }
                ore[((size_t)kz * P + ky) * H + kx] = sr;
                oim[((size_t)kz * P + ky) * H + kx] = si;
            }
    free(are);
    free(aim);
    free(bre);
    free(bim);
}

/* c2r restricted to the output window [n-1, 2n-2]^3; out += scale * result / P^3 */
static void dft_inverse_window(const dft_t* d, const double* cre, const double* cim, double scale, double* out)
{
    const int n = d->n, P = d->P, H = d->H;
    size_t    sd = (size_t)n * P * H, se = (size_t)n * n * H;
    double *  dre = (double*)malloc(sizeof(double) * sd), *dim = (double*)malloc(sizeof(double) * sd);
    double *  ere = (double*)malloc(sizeof(double) * se), *eim = (double*)malloc(sizeof(double) * se);
    const double f = scale / ((double)P * P * P);
    int       x, y, z, kx, ky, kz;
    for (z = 0; z < n; ++z)
        for (ky = 0; ky < P; ++ky)
            for (kx = 0; kx < H; ++kx)
            {
                double sr = 0.0, si = 0.0;
                for (kz = 0; kz < P; ++kz)
                {
                    int    m = (kz * (n - 1 + z)) % P;
                    double c = d->cw[m], s = d->sw[m];
                    double ar = cre[((size_t)kz * P + ky) * H + kx], ai = cim[((size_t)kz * P + ky) * H + kx];
                    sr += ar * c - ai * s;
                    si += ar * s + ai * c;
                }
                dre[((size_t)z * P + ky) * H + kx] = sr;
                dim[((size_t)z * P + ky) * H + kx] = si;
            }
    for (z = 0; z < n; ++z)
        for (y = 0; y < n; ++y)
            for (kx = 0; kx < H; ++kx)
            {
                double sr = 0.0, si = 0.0;
                for (ky = 0; ky < P; ++ky)
                {
                    int    m = (ky * (n - 1 + y)) % P;
                    double c = d->cw[m], s = d->sw[m];
                    double ar = dre[((size_t)z * P + ky) * H + kx], ai = dim[((size_t)z * P + ky) * H + kx];
                    sr += ar * c - ai * s;
                    si += ar * s + ai * c;
                }
                ere[((size_t)z * n + y) * H + kx] = sr;
                eim[((size_t)z * n + y) * H + kx] = si;
            }
    for (z = 0; z < n; ++z)
        for (y = 0; y < n; ++y)
            for (x = 0; x < n; ++x)
            {
                double s = 0.0;
                for (kx = 0; kx < H; ++kx)
                {
                    int    m = (kx * (n - 1 + x)) % P;
                    double w = (kx == 0 || (P % 2 == 0 && 2 * kx == P)) ? 1.0 : 2.0;
                    double ar = ere[((size_t)z * n + y) * H + kx], ai = eim[((size_t)z * n + y) * H + kx];
                    s += w * (ar * d->cw[m] - ai * d->sw[m]);
                }
                out[((size_t)z * n + y) * n + x] += f * s;
            }
    free(dre);
    free(dim);
    free(ere);
    free(eim);
}

/* ------------------------------------------------------------------------ */
/* Kernels: Laplace LGF or integrating factor                                 */
/* ------------------------------------------------------------------------ */
typedef struct
{
    int           is_if;
    const double* g;
    int           nmax;
    double        alpha; /* current level alpha (IF) */
} kernel_t;

static double kernel_value(const kernel_t* k, int x, int y, int z, double* ivbuf, int ivmax)
{
    if (!k->is_if) return oracle_lgf_get(k->g, k->nmax, x, y, z);
    (void)ivbuf;
    (void)ivmax;
    {
        int     m = abs(x), t;
        double* iv;
        double  r;
        if (abs(y) > m) m = abs(y);
        if (abs(z) > m) m = abs(z);
        iv = (double*)malloc(sizeof(double) * (m + 1));
        oracle_ive(2.0 * k->alpha, m, iv);
        r = iv[abs(x)] * iv[abs(y)] * iv[abs(z)];
        (void)t;
        free(iv);
        return r;
    }
}

/* Spectrum cache keyed by (level_diff, lgf block base) as LGF_Base::dft. */
typedef struct
{
    int     d, bx, by, bz;
    double *re, *im;
} spec_entry;
typedef struct
{
    spec_entry* e;
    int         n, cap;
} spec_cache;

static void cache_clear(spec_cache* c)
{
    int i;
    for (i = 0; i < c->n; ++i)
    {
        free(c->e[i].re);
        free(c->e[i].im);
    }
    c->n = 0;
}

static spec_entry* cache_find(spec_cache* c, int ld, const int base[3])
{
    int i;
    for (i = 0; i < c->n; ++i)
        if (c->e[i].d == ld && c->e[i].bx == base[0] && c->e[i].by == base[1] && c->e[i].bz == base[2])
            return &c->e[i];
    return NULL;
}

/* a new entry with no spectrum yet */
static spec_entry* cache_add(spec_cache* c, int ld, const int base[3])
{
    spec_entry* s;
    if (c->n == c->cap)
    {
        c->cap = c->cap ? 2 * c->cap : 64;
        c->e = (spec_entry*)realloc(c->e, sizeof(spec_entry) * c->cap);
    }
    s = &c->e[c->n++];
    s->d = ld;
    s->bx = base[0];
    s->by = base[1];
    s->bz = base[2];
    s->re = s->im = NULL;
    return s;
}

/* LGF_Base::dft (lgf.hpp:77-98, 221-262): the kernel sampled at step 2^d on
 * the (2n-1)^3 LGF block, then r2c */
static void spec_fill(spec_entry* s, const dft_t* d, const kernel_t* k)
{
    const int P = d->P, step = 1 << s->d;
    size_t    P3 = (size_t)P * P * P;
    double*   blk = (double*)malloc(sizeof(double) * P3);
    int       x, y, z;
    double*   re = (double*)malloc(sizeof(double) * spec_size(d));
    double*   im = (double*)malloc(sizeof(double) * spec_size(d));
    for (z = 0; z < P; ++z)
        for (y = 0; y < P; ++y)
            for (x = 0; x < P; ++x)
                blk[((size_t)z * P + y) * P + x] = kernel_value(
                    k, (s->bx + x) * step, (s->by + y) * step, (s->bz + z) * step, NULL, 0);
    dft_forward(d, blk, P, re, im);
    free(blk);
    s->re = re;
    s->im = im;
}

static const spec_entry* lgf_spectrum(spec_cache* c, const dft_t* d, const kernel_t* k, int ld,
    const int base[3])
{
    spec_entry* s = cache_find(c, ld, base);
    if (s && s->re) return s;
    if (!s) s = cache_add(c, ld, base);
    spec_fill(s, d, k);
    return s;
}

void oracle_block_conv(const double* g, int nmax, int n, const double* src, const int shift[3],
    int step, double scale, double* tgt)
{
    dft_t      d;
    spec_cache c = {0, 0, 0};
    kernel_t   k = {0, g, nmax, 0.0};
    int        ld = 0, base[3], i;
    size_t     S;
    double *   sre, *sim, *are, *aim;
    const spec_entry* e;
    while ((1 << ld) < step) ++ld;
    dft_init(&d, n);
    S = spec_size(&d);
    for (i = 0; i < 3; ++i) base[i] = shift[i] - (n - 1);
    e = lgf_spectrum(&c, &d, &k, ld, base);
    sre = (double*)malloc(sizeof(double) * S);
    sim = (double*)malloc(sizeof(double) * S);
    are = (double*)calloc(S, sizeof(double));
    aim = (double*)calloc(S, sizeof(double));
    dft_forward(&d, src, n, sre, sim);
    for (i = 0; i < (int)S; ++i)
    {
        are[i] += e->re[i] * sre[i] - e->im[i] * sim[i];
        aim[i] += e->re[i] * sim[i] + e->im[i] * sre[i];
    }
    dft_inverse_window(&d, are, aim, scale, tgt);
    free(sre);
    free(sim);
    free(are);
    free(aim);
    cache_clear(&c);
    free(c.e);
    dft_free(&d);
}

/* ------------------------------------------------------------------------ */
/* FMM and Poisson solver                                                     */
/* ------------------------------------------------------------------------ */
typedef struct
{
    const oracle_tree* t;
    int                n;
    size_t             n3;
    dft_t              dft;
    spec_cache         cache;
    kernel_t           ker;
    int                faithful;
    int                intrp_order;
    int                ranks; /* > 0: Bx and source transforms on `ranks` workers, contiguous load split */
    /* transfer matrices: nli (anterp/interp) and cell-centre (restrict/interp) */
    double *nli_sub[2], *nli_subT[2], *cc_restrict[2], *cc_interpT[2];
    /* the same at relative position 0 (samples coincide with the child's) */
    double *cc_restrict0[2], *cc_interpT0[2];
    /* mesh type of the solved field (0 cell, 1 face, 2 edge) and its field
     * index: per axis, the relative position of the level transfers
     * (cell_center_nli_intrp.hpp:599-618) */
    int mesh, comp;
    /* working fields [nb][n^3] */
    double *source_tmp, *target_tmp, *correction_tmp, *fmm_s, *fmm_t;
    /* per-pass source spectra (non-faithful mode) */
    double **src_re, **src_im;
    int*    lvl_begin; /* slots of level l: [lvl_begin[l], lvl_begin[l+1]) */
} ostate;

#define BLK(f, s) ((f) + (size_t)(s) * o->n3)
#define IS_LEAF(s) (o->t->flags[s] & 1u)
#define IS_CORR(s) (o->t->flags[s] & 2u)
#define MS(s, idx) ((o->t->mask_src[s] >> (idx)) & 1ull)
#define MT(s, idx) ((o->t->mask_tgt[s] >> (idx)) & 1ull)

static void ostate_init(ostate* o, const oracle_tree* t, int faithful, int intrp_order)
{
    int     n = t->n, b, l;
    double* A = (double*)malloc(sizeof(double) * n * 2 * n);
    double* S = (double*)malloc(sizeof(double) * n * n);
    memset(o, 0, sizeof(*o));
    o->t = t;
    o->n = n;
    o->n3 = (size_t)n * n * n;
    o->faithful = faithful;
    o->intrp_order = intrp_order;
    dft_init(&o->dft, n);
    oracle_lagrange_matrix(n, 10, 0, A);
    for (b = 0; b < 2; ++b)
    {
        o->nli_sub[b] = (double*)malloc(sizeof(double) * n * n);
        o->nli_subT[b] = (double*)malloc(sizeof(double) * n * n);
        sub_matrix(A, n, b, o->nli_sub[b]);
        transpose(o->nli_sub[b], n, o->nli_subT[b]);
    }
    oracle_lagrange_matrix(n, intrp_order, 1, A);
    for (b = 0; b < 2; ++b)
    {
        o->cc_restrict[b] = (double*)malloc(sizeof(double) * n * n);
        o->cc_interpT[b] = (double*)malloc(sizeof(double) * n * n);
        restrict_matrix(n, b, 1, o->cc_restrict[b]);
        sub_matrix(A, n, b, S);
        transpose(S, n, o->cc_interpT[b]);
    }
    oracle_lagrange_matrix(n, intrp_order, 0, A);
    for (b = 0; b < 2; ++b)
    {
        o->cc_restrict0[b] = (double*)malloc(sizeof(double) * n * n);
        o->cc_interpT0[b] = (double*)malloc(sizeof(double) * n * n);
        restrict_matrix(n, b, 0, o->cc_restrict0[b]);
        sub_matrix(A, n, b, S);
        transpose(S, n, o->cc_interpT0[b]);
    }
    free(A);
    free(S);
    o->source_tmp = (double*)calloc(o->n3 * t->nb, sizeof(double));
    o->target_tmp = (double*)calloc(o->n3 * t->nb, sizeof(double));
    o->correction_tmp = (double*)calloc(o->n3 * t->nb, sizeof(double));
    o->fmm_s = (double*)calloc(o->n3 * t->nb, sizeof(double));
    o->fmm_t = (double*)calloc(o->n3 * t->nb, sizeof(double));
    o->src_re = (double**)calloc(t->nb, sizeof(double*));
    o->src_im = (double**)calloc(t->nb, sizeof(double*));
    o->lvl_begin = (int*)calloc(t->depth + 2, sizeof(int));
    for (l = 0; l <= t->depth; ++l)
    {
        int s = 0;
        while (s < t->nb && t->level[s] < l) ++s;
        o->lvl_begin[l] = s;
    }
}

static void ostate_free(ostate* o)
{
    int b, s;
    for (b = 0; b < 2; ++b)
    {
        free(o->nli_sub[b]);
        free(o->nli_subT[b]);
        free(o->cc_restrict[b]);
        free(o->cc_interpT[b]);
        free(o->cc_restrict0[b]);
        free(o->cc_interpT0[b]);
    }
    for (s = 0; s < o->t->nb; ++s)
    {
        free(o->src_re[s]);
        free(o->src_im[s]);
    }
    free(o->src_re);
    free(o->src_im);
    free(o->source_tmp);
    free(o->target_tmp);
    free(o->correction_tmp);
    free(o->fmm_s);
    free(o->fmm_t);
    free(o->lvl_begin);
    cache_clear(&o->cache);
    free(o->cache.e);
    dft_free(&o->dft);
}

static void child_bits(int c, int* bx, int* by, int* bz)
{
    *bx = c & 1;
    *by = (c >> 1) & 1;
    *bz = (c >> 2) & 1;
}

/* fmm_tt + execute_fwrd_field: accumulate G_hat * S_hat into (are, aim) */
static void fmm_tt(ostate* o, int s, int tg, int ld, double* are, double* aim)
{
    const int*  bb = o->t->bbase;
    int         base[3], i;
    size_t      S = spec_size(&o->dft), k;
    const spec_entry* e;
    double *    sre, *sim;
    int         own = 0;
    for (i = 0; i < 3; ++i) base[i] = (bb[3 * tg + i] - bb[3 * s + i]) - (o->n - 1);
    e = lgf_spectrum(&o->cache, &o->dft, &o->ker, ld, base);
    if (o->faithful || !o->src_re[s])
    {
        sre = (double*)malloc(sizeof(double) * S);
        sim = (double*)malloc(sizeof(double) * S);
        dft_forward(&o->dft, BLK(o->fmm_s, s), o->n, sre, sim);
        if (o->faithful) own = 1;
        else
        {
            o->src_re[s] = sre;
            o->src_im[s] = sim;
        }
    }
    else
    {
        sre = o->src_re[s];
        sim = o->src_im[s];
    }
    for (k = 0; k < S; ++k)
    {
        are[k] += e->re[k] * sre[k] - e->im[k] * sim[k];
        aim[k] += e->re[k] * sim[k] + e->im[k] * sre[k];
    }
    if (own)
    {
        free(sre);
        free(sim);
    }
}

/* compute_influence_field (fmm.hpp:988-1027) for one target */
static long influence_field(ostate* o, int tg, int L, int idx, int neighbor_only, double scale)
{
    const oracle_tree* t = o->t;
    size_t             S = spec_size(&o->dft);
    double *           are = (double*)calloc(S, sizeof(double)), *aim = (double*)calloc(S, sizeof(double));
    long               count = 0;
    int                i;
    const int          lvl = t->level[tg];
    if (lvl == L)
        for (i = 0; i < 27; ++i)
        {
            int s = t->neighbor[27 * tg + i];
            if (s >= 0 && MS(s, idx))
            {
                fmm_tt(o, s, tg, 0, are, aim);
                ++count;
            }
        }
    if (!neighbor_only)
        for (i = 0; i < t->influence_num[tg]; ++i)
        {
            int s = t->influence[189 * tg + i];
            if (s >= 0 && MS(s, idx))
            {
                fmm_tt(o, s, tg, L - lvl, are, aim);
                ++count;
            }
        }
    if (count) dft_inverse_window(&o->dft, are, aim, scale, BLK(o->fmm_t, tg));
    free(are);
    free(aim);
    return count;
}

/* Server::split (include/iblgf/domain/decomposition/server.hpp:127-170): items
 * in their tree order go to W workers as contiguous ranges, worker w taking
 * items while the running load is at most (w + 1)/W of the total (and at least
 * one item); a zero load counts as 1.  begin[w]..begin[w+1] is worker w's range. */
static void split_ranges(const long* load, int n, int W, int* begin)
{
    double total = 0.0, current = 0.0;
    int    i, w, it = 0;
    for (i = 0; i < n; ++i) total += load[i] > 0 ? (double)load[i] : 1.0;
    for (w = 0; w <= W; ++w) begin[w] = n;
    for (w = 0; w < W && it < n; ++w)
    {
        double target = (double)(w + 1) / W * total;
        int    count = 0;
        if (target > total) target = total;
        begin[w] = it;
        while (it < n && (current <= target || count == 0))
        {
            current += load[it] > 0 ? (double)load[it] : 1.0;
            ++it;
            ++count;
        }
    }
    for (; w <= W; ++w) begin[w] = n;
}

/* pairs a target takes in one pass: its load for the split */
static long target_load(const ostate* o, int tg, int L, int idx, int neighbor_only)
{
    const oracle_tree* t = o->t;
    long               c = 0;
    int                i;
    if (t->level[tg] == L)
        for (i = 0; i < 27; ++i)
        {
            const int s = t->neighbor[27 * tg + i];
            if (s >= 0 && MS(s, idx)) ++c;
        }
    if (!neighbor_only)
        for (i = 0; i < t->influence_num[tg]; ++i)
        {
            const int s = t->influence[189 * tg + i];
            if (s >= 0 && MS(s, idx)) ++c;
        }
    return c;
}

static void source_spectrum(ostate* o, int s)
{
    const size_t S = spec_size(&o->dft);
    double*      re = (double*)malloc(sizeof(double) * S);
    double*      im = (double*)malloc(sizeof(double) * S);
    dft_forward(&o->dft, BLK(o->fmm_s, s), o->n, re, im);
    o->src_re[s] = re;
    o->src_im[s] = im;
}

/* Every kernel spectrum (and, when cached, every source spectrum) that the
 * targets tg[0..nt) will read, made ahead of the concurrent Bx loop. */
static void prepare_spectra(ostate* o, const int* tg, int nt, int L, int idx, int neighbor_only)
{
    const oracle_tree* t = o->t;
    int                i, j, first = o->cache.n, nnew, src;
    char*              need = (char*)calloc(t->nb, 1);
    int*               list = (int*)malloc(sizeof(int) * (size_t)(t->nb + 1));
    int                nl = 0;
    for (i = 0; i < nt; ++i)
    {
        const int s = tg[i], lvl = t->level[s];
        for (j = 0; j < 27 + t->influence_num[s]; ++j)
        {
            int base[3], d;
            if (j < 27) src = lvl == L ? t->neighbor[27 * s + j] : -1;
            else src = neighbor_only ? -1 : t->influence[189 * s + j - 27];
            if (src < 0 || !MS(src, idx)) continue;
            for (d = 0; d < 3; ++d) base[d] = (t->bbase[3 * s + d] - t->bbase[3 * src + d]) - (o->n - 1);
            if (!cache_find(&o->cache, j < 27 ? 0 : L - lvl, base)) cache_add(&o->cache, j < 27 ? 0 : L - lvl, base);
            if (!o->faithful && !o->src_re[src] && !need[src])
            {
                need[src] = 1;
                list[nl++] = src;
            }
        }
    }
    nnew = o->cache.n - first;
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < nnew; ++i) spec_fill(&o->cache.e[first + i], &o->dft, &o->ker);
    if (o->ranks > 0)
    {
        /* the sources in contiguous equal shares, one per worker */
        int*  b = (int*)malloc(sizeof(int) * (size_t)(o->ranks + 1));
        long* ld = (long*)malloc(sizeof(long) * (size_t)(nl + 1));
        for (i = 0; i < nl; ++i) ld[i] = 1;
        split_ranges(ld, nl, o->ranks, b);
#pragma omp parallel num_threads(o->ranks)
        {
            int w, k;
            for (w = worker_id(); w < o->ranks; w += worker_count())
                for (k = b[w]; k < b[w + 1]; ++k) source_spectrum(o, list[k]);
        }
        free(b);
        free(ld);
    }
    else
    {
#pragma omp parallel for schedule(dynamic, 1)
        for (i = 0; i < nl; ++i) source_spectrum(o, list[i]);
    }
    free(need);
    free(list);
}

/* Fmm::apply (fmm.hpp:729-870) with fmm mask index idx */
static void fmm_apply_idx(ostate* o, const double* Source, double* Target, int L, int idx, double add_scale)
{
    const oracle_tree* t = o->t;
    const int          rl = L - t->base_level;
    const int          neighbor_only = o->ker.is_if;
    const double       dx_level = t->dx_base / pow(2.0, rl);
    int                l, s, c, i, j, k;
    const int          n = o->n;
    /* init zero (levels L..0) */
    for (l = L; l >= 0; --l)
        for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
        {
            if (MS(s, idx)) memset(BLK(o->fmm_s, s), 0, sizeof(double) * o->n3);
            if (MT(s, idx)) memset(BLK(o->fmm_t, s), 0, sizeof(double) * o->n3);
        }
    /* init copy: interior only */
    for (s = o->lvl_begin[L]; s < o->lvl_begin[L + 1]; ++s)
        if (MS(s, idx))
            for (k = 1; k < n - 1; ++k)
                for (j = 1; j < n - 1; ++j)
                    for (i = 1; i < n - 1; ++i)
                    {
                        size_t q = ((size_t)k * n + j) * n + i;
                        BLK(o->fmm_s, s)[q] = BLK(Source, s)[q];
                    }
    for (s = 0; s < t->nb; ++s)
    {
        free(o->src_re[s]);
        free(o->src_im[s]);
        o->src_re[s] = o->src_im[s] = NULL;
    }
    if (!neighbor_only)
    {
        /* anterpolation (fmm.hpp:1128-1165) */
        for (l = L - 1; l >= 0; --l)
            for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
            {
                if (!MS(s, idx)) continue;
                for (c = 0; c < 8; ++c)
                {
                    int ch = t->children[8 * s + c], bx, by, bz;
                    if (ch < 0 || !MS(ch, idx)) continue;
                    child_bits(c, &bx, &by, &bz);
                    transfer3(n, o->nli_sub[bx], o->nli_sub[by], o->nli_sub[bz], BLK(o->fmm_s, ch),
                        BLK(o->fmm_s, s), 1);
                }
            }
    }
    /* Bx over targets at levels L..0 (L only for neighbour-only kernels).  The
     * kernel and source spectra the targets read are made first; then targets
     * run concurrently, each writing only its own fmm_t block and summing its
     * sources in the serial order, so results do not depend on the threads. */
    {
        int* tg = (int*)malloc(sizeof(int) * (size_t)(t->nb + 1));
        int  nt = 0, it;
        for (l = L; l >= (neighbor_only ? L : 0); --l)
            for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
                if (MT(s, idx)) tg[nt++] = s;
        prepare_spectra(o, tg, nt, L, idx, neighbor_only);
        if (o->ranks > 0)
        {
            /* the targets, in tree order, as contiguous ranges of equal pair load */
            int*  b = (int*)malloc(sizeof(int) * (size_t)(o->ranks + 1));
            long* ld = (long*)malloc(sizeof(long) * (size_t)(nt + 1));
            for (it = 0; it < nt; ++it) ld[it] = target_load(o, tg[it], L, idx, neighbor_only);
            split_ranges(ld, nt, o->ranks, b);
#pragma omp parallel num_threads(o->ranks)
            {
                int w, k;
                for (w = worker_id(); w < o->ranks; w += worker_count())
                    for (k = b[w]; k < b[w + 1]; ++k)
                        influence_field(o, tg[k], L, idx, neighbor_only, neighbor_only ? 1.0 : dx_level * dx_level);
            }
            free(b);
            free(ld);
        }
        else
        {
#pragma omp parallel for schedule(dynamic, 1)
            for (it = 0; it < nt; ++it)
                influence_field(o, tg[it], L, idx, neighbor_only, neighbor_only ? 1.0 : dx_level * dx_level);
        }
        free(tg);
    }
    if (!neighbor_only)
    {
        /* interpolation (fmm.hpp:1108-1126) */
        for (l = 1; l < L; ++l)
            for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
            {
                if (!MT(s, idx)) continue;
                for (c = 0; c < 8; ++c)
                {
                    int ch = t->children[8 * s + c], bx, by, bz;
                    if (ch < 0 || !MT(ch, idx)) continue;
                    child_bits(c, &bx, &by, &bz);
                    transfer3(n, o->nli_subT[bx], o->nli_subT[by], o->nli_subT[bz], BLK(o->fmm_t, s),
                        BLK(o->fmm_t, ch), 1);
                }
            }
    }
    /* fmm_add_equal */
    for (s = o->lvl_begin[L]; s < o->lvl_begin[L + 1]; ++s)
        if (MT(s, idx))
        {
            size_t q;
            for (q = 0; q < o->n3; ++q) BLK(Target, s)[q] += BLK(o->fmm_t, s)[q] * add_scale;
        }
}

/* AMR2AMR mask index 15 + 2 rl + nl (octant.hpp:94-131) */
static void fmm_apply(ostate* o, const double* Source, double* Target, int L, int nl, double add_scale)
{
    fmm_apply_idx(o, Source, Target, L, 15 + 2 * (L - o->t->base_level) + nl, add_scale);
}

/* halo exchange at one level: buffer cells <- neighbour interiors */
static void buffer_exchange(ostate* o, double* f, int l)
{
    const oracle_tree* t = o->t;
    const int          n = o->n, e = n - 2;
    int                s, i, j, k;
    for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
        for (k = 0; k < n; ++k)
            for (j = 0; j < n; ++j)
                for (i = 0; i < n; ++i)
                {
                    int dx = (i == 0) ? -1 : (i == n - 1 ? 1 : 0);
                    int dy = (j == 0) ? -1 : (j == n - 1 ? 1 : 0);
                    int dz = (k == 0) ? -1 : (k == n - 1 ? 1 : 0);
                    int nb;
                    if (!dx && !dy && !dz) continue;
                    nb = t->neighbor[27 * s + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
                    if (nb < 0) continue;
                    BLK(f, s)[((size_t)k * n + j) * n + i] =
                        BLK(f, nb)[((size_t)(k - dz * e) * n + (j - dy * e)) * n + (i - dx * e)];
                }
}

/* relative position (0 or 1) of the solved field's samples on one axis */
static int rel_pos(const ostate* o, int axis)
{
    if (o->mesh == 1) return axis == o->comp ? 0 : 1; /* face */
    if (o->mesh == 2) return axis == o->comp ? 1 : 0; /* edge */
    return 1;                                          /* cell */
}
static const double* restrict_of(const ostate* o, int axis, int bit)
{
    return rel_pos(o, axis) ? o->cc_restrict[bit] : o->cc_restrict0[bit];
}
static const double* interp_of(const ostate* o, int axis, int bit)
{
    return rel_pos(o, axis) ? o->cc_interpT[bit] : o->cc_interpT0[bit];
}

/* cell-centred restriction of all children of octants at level ls into the parent */
static void source_coarsify(ostate* o, double* f)
{
    const oracle_tree* t = o->t;
    int                ls, s, c;
    for (ls = t->depth - 2; ls >= t->base_level; --ls)
        for (s = o->lvl_begin[ls]; s < o->lvl_begin[ls + 1]; ++s)
            for (c = 0; c < 8; ++c)
            {
                int ch = t->children[8 * s + c], bx, by, bz;
                if (ch < 0) continue;
                child_bits(c, &bx, &by, &bz);
                transfer3(o->n, restrict_of(o, 0, bx), restrict_of(o, 1, by), restrict_of(o, 2, bz), BLK(f, ch),
                    BLK(f, s), 1);
            }
}

/* cell-centred interpolation parent -> children (assignment) */
static void cc_intrp_node(ostate* o, const double* from, double* to, int s, int correction_only)
{
    const oracle_tree* t = o->t;
    int                c;
    for (c = 0; c < 8; ++c)
    {
        int ch = t->children[8 * s + c], bx, by, bz;
        if (ch < 0) continue;
        if (correction_only && !IS_CORR(ch)) continue;
        child_bits(c, &bx, &by, &bz);
        transfer3(o->n, interp_of(o, 0, bx), interp_of(o, 1, by), interp_of(o, 2, bz), BLK(from, s), BLK(to, ch), 0);
    }
}

/* add_source_correction (cell_center_nli_intrp.hpp:171-238) */
static void add_source_correction(ostate* o, int s, double dx, double omega)
{
    const oracle_tree* t = o->t;
    const int          n = o->n;
    int                c, i, j, k;
    const double       inv = 1.0 / (dx * dx);
    for (c = 0; c < 8; ++c)
    {
        int           ch = t->children[8 * s + c];
        const double* tt;
        double*       cc;
        if (ch < 0) continue;
        tt = BLK(o->target_tmp, ch);
        cc = BLK(o->correction_tmp, ch);
        for (k = 1; k < n - 1; ++k)
            for (j = 1; j < n - 1; ++j)
                for (i = 1; i < n - 1; ++i)
                {
                    size_t q = ((size_t)k * n + j) * n + i;
                    cc[q] += 6.0 * tt[q] * inv;
                    cc[q] -= tt[q - (size_t)n * n] * inv;
                    cc[q] -= tt[q + (size_t)n * n] * inv;
                    cc[q] -= tt[q - n] * inv;
                    cc[q] -= tt[q + n] * inv;
                    cc[q] -= tt[q + 1] * inv;
                    cc[q] -= tt[q - 1] * inv;
                    cc[q] += omega * omega * tt[q];
                }
    }
}

/* apply_lgf's level loop (poisson.hpp:587-894) on an initialised state */
static void lgf_solve(ostate* o, const double* src, double* dst)
{
    const oracle_tree* t = o->t;
    int                l, s;
    size_t             q;
    /* copy_leaf with buffer */
    for (s = 0; s < t->nb; ++s)
        if (IS_LEAF(s)) memcpy(BLK(o->source_tmp, s), BLK(src, s), sizeof(double) * o->n3);
    source_coarsify(o, o->source_tmp);
    for (l = t->base_level; l < t->depth; ++l)
    {
        const double dx = t->dx_base / pow(2.0, l - t->base_level);
        fmm_apply(o, o->source_tmp, o->target_tmp, l, 1, 1.0);
        for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
            cc_intrp_node(o, o->target_tmp, o->target_tmp, s, 0);
        fmm_apply(o, o->source_tmp, o->target_tmp, l, 0, 1.0);
        for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
            if (IS_LEAF(s)) memcpy(BLK(dst, s), BLK(o->target_tmp, s), sizeof(double) * o->n3);
        if (l == t->depth - 1) continue;
        buffer_exchange(o, o->source_tmp, l);
        for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
            cc_intrp_node(o, o->source_tmp, o->correction_tmp, s, 1);
        for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s) add_source_correction(o, s, dx / 2.0, 0.0);
        for (s = o->lvl_begin[l + 1]; s < o->lvl_begin[l + 2]; ++s)
            for (q = 0; q < o->n3; ++q) BLK(o->source_tmp, s)[q] += BLK(o->correction_tmp, s)[q];
    }
}

int oracle_apply_lgf_mesh(const oracle_tree* t, const double* g, int nmax, const double* src, double* dst, int mesh,
    int comp)
{
    ostate  o_;
    ostate* o = &o_;
    if (mesh < 0 || mesh > 2 || comp < 0 || comp > 2) return 1;
    ostate_init(o, t, 0, 5);
    o->mesh = mesh;
    o->comp = comp;
    o->ker.is_if = 0;
    o->ker.g = g;
    o->ker.nmax = nmax;
    lgf_solve(o, src, dst);
    ostate_free(o);
    return 0;
}

int oracle_apply_lgf(const oracle_tree* t, const double* g, int nmax, const double* src, double* dst,
    int faithful, int intrp_order)
{
    ostate  o_;
    ostate* o = &o_;
    ostate_init(o, t, faithful, intrp_order);
    o->ker.is_if = 0;
    o->ker.g = g;
    o->ker.nmax = nmax;
    lgf_solve(o, src, dst);
    ostate_free(o);
    return 0;
}

int oracle_lgf_baseline(const oracle_tree* t, const double* g, int nmax, const double* src, double* dst,
    int faithful, int ranks, int nsolves, double* seconds)
{
    ostate  o_;
    ostate* o = &o_;
    int     i;
    ostate_init(o, t, faithful, 5);
    o->ker.is_if = 0;
    o->ker.g = g;
    o->ker.nmax = nmax;
    o->ranks = ranks;
    for (i = 0; i <= nsolves; ++i)
    {
        const size_t all = o->n3 * (size_t)t->nb * sizeof(double);
        const double t0 = wall_seconds();
        /* every solve starts from cleared work fields, as clean_field does
         * (poisson.hpp:597-600); the kernel spectra stay cached */
        memset(o->source_tmp, 0, all);
        memset(o->target_tmp, 0, all);
        memset(o->correction_tmp, 0, all);
        memset(o->fmm_s, 0, all);
        memset(o->fmm_t, 0, all);
        lgf_solve(o, src, dst);
        if (i > 0) seconds[i - 1] = wall_seconds() - t0;
    }
    ostate_free(o);
    return 0;
}

static int if_solve(const oracle_tree* t, const double* src, double* dst, double alpha_base, int faithful,
    int intrp_order, int mesh, int comp);

int oracle_apply_if(const oracle_tree* t, const double* src, double* dst, double alpha_base, int faithful,
    int intrp_order)
{
    return if_solve(t, src, dst, alpha_base, faithful, intrp_order, 0, 0);
}

int oracle_apply_if_mesh(const oracle_tree* t, const double* src, double* dst, double alpha_base, int mesh, int comp)
{
    if (mesh < 0 || mesh > 2 || comp < 0 || comp > 2) return 1;
    return if_solve(t, src, dst, alpha_base, 0, 5, mesh, comp);
}

int oracle_mesh_transfer(int n, int mesh, int comp, int op, int child, const double* in, double* out)
{
    oracle_tree t;
    ostate      o_;
    ostate*     o = &o_;
    if (n < 4 || mesh < 0 || mesh > 2 || comp < 0 || comp > 2 || op < 0 || op > 1 || child < 0 || child > 7)
        return 1;
    memset(&t, 0, sizeof(t));
    t.n = n; /* no blocks: only the transfer matrices are built */
    ostate_init(o, &t, 0, 5);
    o->mesh = mesh;
    o->comp = comp;
    {
        int bx, by, bz;
        child_bits(child, &bx, &by, &bz);
        if (op == 0)
            transfer3(n, restrict_of(o, 0, bx), restrict_of(o, 1, by), restrict_of(o, 2, bz), in, out, 1);
        else
            transfer3(n, interp_of(o, 0, bx), interp_of(o, 1, by), interp_of(o, 2, bz), in, out, 0);
    }
    ostate_free(o);
    return 0;
}

/* apply_if (poisson.hpp:435-494) */
static int if_solve(const oracle_tree* t, const double* src, double* dst, double alpha_base, int faithful,
    int intrp_order, int mesh, int comp)
{
    ostate o_;
    ostate* o = &o_;
    int     l, s;
    ostate_init(o, t, faithful, intrp_order);
    o->mesh = mesh;
    o->comp = comp;
    o->ker.is_if = 1;
    for (s = 0; s < t->nb; ++s)
        if (IS_LEAF(s)) memcpy(BLK(o->source_tmp, s), BLK(src, s), sizeof(double) * o->n3);
    source_coarsify(o, o->source_tmp);
    for (l = t->depth - 2; l >= t->base_level; --l)
    {
        buffer_exchange(o, o->source_tmp, l);
        for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
            cc_intrp_node(o, o->source_tmp, o->source_tmp, s, 1);
    }
    for (l = t->base_level; l < t->depth; ++l)
    {
        o->ker.alpha = alpha_base * pow(4.0, l - t->base_level);
        cache_clear(&o->cache);
        fmm_apply(o, o->source_tmp, o->target_tmp, l, 0, 1.0);
        fmm_apply(o, o->source_tmp, o->target_tmp, l, 1, 1.0);
        for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
            memcpy(BLK(dst, s), BLK(o->target_tmp, s), sizeof(double) * o->n3);
    }
    ostate_free(o);
    return 0;
}

long oracle_bx_sample(const oracle_tree* t, const double* g, int nmax, const double* fmm_s, int L, int nl,
    int ntargets, const int* targets, int faithful, double* fmm_t, double* seconds)
{
    ostate o_;
    ostate* o = &o_;
    long    pairs = 0;
    int     i, j;
    const int rl = L - t->base_level, idx = 15 + 2 * rl + nl;
    const double dx_level = t->dx_base / pow(2.0, rl);
    double  t0;
    ostate_init(o, t, faithful, 5);
    o->ker.is_if = 0;
    o->ker.g = g;
    o->ker.nmax = nmax;
    memcpy(o->fmm_s, fmm_s, sizeof(double) * o->n3 * t->nb);
    /* warm the kernel-spectrum cache first: the timed part is a warm solve */
    for (i = 0; i < ntargets; ++i)
    {
        int s = targets[i], lvl = t->level[targets[i]];
        for (j = 0; j < 27 + t->influence_num[s]; ++j)
        {
            int src = j < 27 ? (lvl == L ? t->neighbor[27 * s + j] : -1) : t->influence[189 * s + j - 27];
            int base[3], d;
            if (src < 0 || !MS(src, idx)) continue;
            for (d = 0; d < 3; ++d) base[d] = (t->bbase[3 * s + d] - t->bbase[3 * src + d]) - (o->n - 1);
            lgf_spectrum(&o->cache, &o->dft, &o->ker, j < 27 ? 0 : L - lvl, base);
        }
    }
    t0 = wall_seconds();
    for (i = 0; i < ntargets; ++i)
    {
        int s = targets[i];
        memset(BLK(o->fmm_t, s), 0, sizeof(double) * o->n3);
        pairs += influence_field(o, s, L, idx, 0, dx_level * dx_level);
        memcpy(fmm_t + (size_t)i * o->n3, BLK(o->fmm_t, s), sizeof(double) * o->n3);
    }
    if (seconds) *seconds = wall_seconds() - t0;
    ostate_free(o);
    return pairs;
}

/* ======================================================================== */
/* Immersed-boundary solves, block operators, IB transfer, force CG          */
/* ======================================================================== */

#define IS_IB(s) (o->t->flags[s] & 16u)

/* apply_lgf with fmm_type IB2xIB / xIB2IB (poisson.hpp:587-894): copy_level of
 * the finest level's interior, no coarsening, one pass at the finest level,
 * copy of the leaves' target_tmp. */
int oracle_apply_lgf_masked(const oracle_tree* t, const double* g, int nmax, const double* src, double* dst,
    int mask_type, int faithful, int intrp_order)
{
    ostate o_;
    ostate* o = &o_;
    /* STREAM (0): the base level (poisson.hpp:609-611, 645-647); the IB types: the finest */
    const int L = mask_type == 0 ? t->base_level : t->depth - 1;
    const int idx = mask_type == 0 ? 0 : (mask_type == 2 ? 1 : 2);
    int       s, i, j, k, n = t->n;
    if (mask_type != 0 && mask_type != 2 && mask_type != 3 && mask_type != 4) return 1;
    ostate_init(o, t, faithful, intrp_order);
    o->ker.is_if = 0;
    o->ker.g = g;
    o->ker.nmax = nmax;
    if (mask_type == 4)
    {
        /* IB2AMR (poisson.hpp:613-616, 632-633, 645-650): the finest level's
         * interiors coarsened to every level, then per level l one pass with
         * the IB2AMR masks of l (mask index 4 + l - base), its leaves written */
        int l;
        for (s = o->lvl_begin[L]; s < o->lvl_begin[L + 1]; ++s)
            for (k = 1; k < n - 1; ++k)
                for (j = 1; j < n - 1; ++j)
                    for (i = 1; i < n - 1; ++i)
                    {
                        size_t q = ((size_t)k * n + j) * n + i;
                        BLK(o->source_tmp, s)[q] = BLK(src, s)[q];
                    }
        source_coarsify(o, o->source_tmp);
        for (l = t->base_level; l < t->depth; ++l)
        {
            fmm_apply_idx(o, o->source_tmp, o->target_tmp, l, 4 + l - t->base_level, 1.0);
            for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s)
                if (IS_LEAF(s)) memcpy(BLK(dst, s), BLK(o->target_tmp, s), sizeof(double) * o->n3);
        }
        ostate_free(o);
        return 0;
    }
    for (s = o->lvl_begin[L]; s < o->lvl_begin[L + 1]; ++s)
        for (k = 1; k < n - 1; ++k)
            for (j = 1; j < n - 1; ++j)
                for (i = 1; i < n - 1; ++i)
                {
                    size_t q = ((size_t)k * n + j) * n + i;
                    BLK(o->source_tmp, s)[q] = BLK(src, s)[q];
                }
    fmm_apply_idx(o, o->source_tmp, o->target_tmp, L, idx, 1.0);
    for (s = o->lvl_begin[L]; s < o->lvl_begin[L + 1]; ++s)
        if (IS_LEAF(s)) memcpy(BLK(dst, s), BLK(o->target_tmp, s), sizeof(double) * o->n3);
    ostate_free(o);
    return 0;
}

/* apply_if with fmm_type IB2xIB / xIB2IB (poisson.hpp:435-494): one
 * neighbour-only pass at the finest level, copy_level with buffers. */
int oracle_apply_if_masked(const oracle_tree* t, const double* src, double* dst, double alpha_base, int mask_type,
    int faithful, int intrp_order)
{
    ostate o_;
    ostate* o = &o_;
    const int L = t->depth - 1;
    int       s, i, j, k, n = t->n;
    ostate_init(o, t, faithful, intrp_order);
    o->ker.is_if = 1;
    o->ker.alpha = alpha_base * pow(4.0, L - t->base_level);
    for (s = o->lvl_begin[L]; s < o->lvl_begin[L + 1]; ++s)
        for (k = 1; k < n - 1; ++k)
            for (j = 1; j < n - 1; ++j)
                for (i = 1; i < n - 1; ++i)
                {
                    size_t q = ((size_t)k * n + j) * n + i;
                    BLK(o->source_tmp, s)[q] = BLK(src, s)[q];
                }
    fmm_apply_idx(o, o->source_tmp, o->target_tmp, L, mask_type == 2 ? 1 : 2, 1.0);
    for (s = o->lvl_begin[L]; s < o->lvl_begin[L + 1]; ++s)
        memcpy(BLK(dst, s), BLK(o->target_tmp, s), sizeof(double) * o->n3);
    ostate_free(o);
    return 0;
}

static int level_first(const oracle_tree* t, int l)
{
    int s = 0;
    while (s < t->nb && t->level[s] < l) ++s;
    return s;
}

/* Operator::laplace / gradient / divergence / curl / curl_transpose
 * (operators.hpp:1387-1511, 2420-2453) on the interior nodes of every block
 * at level l.  Fields are [nb][n^3], x fastest; vector fields 3 arrays. */
void oracle_block_op(const oracle_tree* t, int op, const double* s0, const double* s1, const double* s2, double* d0,
    double* d1, double* d2, int l, double scale)
{
    const int    n = t->n;
    const size_t n3 = (size_t)n * n * n, X = 1, Y = n, Z = (size_t)n * n;
    const double dx = t->dx_base / pow(2.0, l - t->base_level);
    int          b, i, j, k;
    for (b = level_first(t, l); b < t->nb && t->level[b] == l; ++b)
        for (k = 1; k < n - 1; ++k)
            for (j = 1; j < n - 1; ++j)
                for (i = 1; i < n - 1; ++i)
                {
                    const size_t q = (size_t)b * n3 + (k * Z + j * Y + i * X);
                    if (op == 0)
                    {
                        const double fac = 1.0 / (dx * dx);
                        d0[q] = -6.0 * s0[q] + s0[q - Z] + s0[q + Z] + s0[q - Y] + s0[q + Y] + s0[q - X] + s0[q + X];
                        d0[q] *= fac;
                    }
                    else if (op == 1)
                    {
                        const double fac = 1.0 / dx;
                        d0[q] = fac * (s0[q] - s0[q - X]);
                        d1[q] = fac * (s0[q] - s0[q - Y]);
                        d2[q] = fac * (s0[q] - s0[q - Z]);
                    }
                    else if (op == 2)
                    {
                        const double fac = 1.0 / dx;
                        d0[q] = -s0[q] - s1[q] - s2[q] + s0[q + X] + s1[q + Y] + s2[q + Z];
                        d0[q] *= fac;
                    }
                    else if (op == 3)
                    {
                        const double fac = 1.0 / dx;
                        d0[q] = s2[q] - s2[q - Y] - s1[q] + s1[q - Z];
                        d0[q] *= fac;
                        d1[q] = s0[q] - s0[q - Z] - s2[q] + s2[q - X];
                        d1[q] *= fac;
                        d2[q] = s1[q] - s1[q - X] - s0[q] + s0[q - Y];
                        d2[q] *= fac;
                    }
                    else
                    {
                        const double fac = 1.0 / dx * scale;
                        d0[q] = +s1[q] - s1[q + Z] + s2[q + Y] - s2[q];
                        d0[q] *= fac;
                        d1[q] = +s2[q] - s2[q + X] + s0[q + Z] - s0[q];
                        d1[q] *= fac;
                        d2[q] = +s0[q] - s0[q + Y] + s1[q + X] - s1[q];
                        d2[q] *= fac;
                    }
                }
}

void oracle_buffer_exchange(const oracle_tree* t, double* f, int l)
{
    const int    n = t->n, e = n - 2;
    const size_t n3 = (size_t)n * n * n;
    int          s, i, j, k;
    for (s = level_first(t, l); s < t->nb && t->level[s] == l; ++s)
        for (k = 0; k < n; ++k)
            for (j = 0; j < n; ++j)
                for (i = 0; i < n; ++i)
                {
                    int dx = (i == 0) ? -1 : (i == n - 1 ? 1 : 0);
                    int dy = (j == 0) ? -1 : (j == n - 1 ? 1 : 0);
                    int dz = (k == 0) ? -1 : (k == n - 1 ? 1 : 0);
                    int nb;
                    if (!dx && !dy && !dz) continue;
                    nb = t->neighbor[27 * s + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
                    if (nb < 0) continue;
                    f[s * n3 + ((size_t)k * n + j) * n + i] =
                        f[nb * n3 + ((size_t)(k - dz * e) * n + (j - dy * e)) * n + (i - dx * e)];
                }
}

/* zero `width` planes on the side of face-neighbour slot i (4,10,12,14,16,22) */
static void zero_side(double* blk, int n, int slot, int width)
{
    int i, j, k;
    for (k = 0; k < n; ++k)
        for (j = 0; j < n; ++j)
            for (i = 0; i < n; ++i)
            {
                int hit = (slot == 4 && k < width) || (slot == 22 && k >= n - width) ||
                          (slot == 10 && j < width) || (slot == 16 && j >= n - width) ||
                          (slot == 12 && i < width) || (slot == 14 && i >= n - width);
                if (hit) blk[((size_t)k * n + j) * n + i] = 0.0;
            }
}

/* kind 0: Operator::clean_leaf_correction_boundary (operators.hpp:329-513);
 * kind 1: Operator::clean_ib_region_boundary (73-187). */
void oracle_clean_boundary(const oracle_tree* t, double* f, int l, int kind, int leaf_only, int width)
{
    static const int faces[6] = {4, 10, 12, 14, 16, 22};
    const int        n = t->n;
    const size_t     n3 = (size_t)n * n * n;
    int              s, b;
    if (kind == 0)
    {
        if (leaf_only)
            for (s = level_first(t, l); s < t->nb && t->level[s] == l; ++s)
                if (t->flags[s] & 2u) memset(f + s * n3, 0, n3 * sizeof(double));
        if (l != t->base_level) return;
    }
    for (s = level_first(t, l); s < t->nb && t->level[s] == l; ++s)
        for (b = 0; b < 6; ++b)
        {
            const int nb = t->neighbor[27 * s + faces[b]];
            int       clean;
            if (kind == 0) clean = nb < 0 || (leaf_only && (t->flags[nb] & 2u));
            else
                clean = nb < 0 || !(t->flags[nb] & 16u);
            if (clean) zero_side(f + s * n3, n, faces[b], width);
        }
}

/* ---- immersed boundary ------------------------------------------------- */

static double ddf1(int kind, double x)
{
    const double r = fabs(x), r2 = r * r, s3 = sqrt(3);
    if (kind == 2)
    {
        if (r > 1.5) return 0;
        if (r <= 0.5) return (1.0 + sqrt(-3. * r2 + 1.0)) / 3.0;
        return (5.0 - 3.0 * r - sqrt(1.0 - 3.0 * (1 - r) * (1 - r))) / 6.0;
    }
    if (kind == 0)
    {
        if (r > 2) return 0;
        if (r <= 1.0)
            return 17.0 / 48.0 + s3 * O_PI / 108.0 + r / 4.0 - r2 / 4.0 +
                   (1.0 - 2.0 * r) / 16 * sqrt(-12.0 * r2 + 12.0 * r + 1.0) - s3 / 12.0 * asin(s3 / 2.0 * (2.0 * r - 1.0));
        return 55.0 / 48.0 - s3 * O_PI / 108.0 - 13.0 * r / 12.0 + r2 / 4.0 +
               (2.0 * r - 3.0) / 48.0 * sqrt(-12.0 * r2 + 36.0 * r - 23.0) + s3 / 36.0 * asin(s3 / 2.0 * (2 * r - 3.0));
    }
    if (r > 2.5) return 0;
    if (r <= 0.5) return 3.0 / 8 + O_PI / 32.0 - r2 / 4;
    if (r <= 1.5) return 1.0 / 4 + (1.0 - r) / 8.0 * sqrt(-2.0 + 8 * r - 4 * r2) - 1.0 / 8 * asin(sqrt(2) * (r - 1));
    /* the reference's 1/16 factor of the arcsine is an integer quotient (0) */
    return 17.0 / 16 - O_PI / 64.0 - 3.0 / 4 * r + r2 / 8 + (r - 2.0) / 16.0 * sqrt(-14.0 + 16 * r - 4 * r2);
}

double oracle_delta_1d(int kind, double x) { return ddf1(kind, x); }

/* IB::ib_block_overlap (ib.hpp:648-678) */
static int ib_overlap(const oracle_tree* t, const oracle_ib* ib, int i, int s, int radius_level)
{
    const int    rl = t->level[s] - t->base_level;
    const int    f = 1 << (ib->ib_level - rl);
    const int    e = t->n - 2;
    double       added = 0;
    int          d;
    const double factor = pow(2, ib->ib_level) / t->dx_base;
    if (radius_level == 2) added += ib->ddf_radius + ib->safety_dis + 10.0;
    else if (radius_level == 1)
        added += ib->ddf_radius + 5.0;
    for (d = 0; d < 3; ++d)
    {
        int base = t->bbase[3 * s + d] * f, ext = e * f;
        ext += (int)(2 * added);
        base -= (int)added;
        if ((base + ext - 1 + 1) < ib->x[3 * i + d] * factor || base >= ib->x[3 * i + d] * factor) return 0;
    }
    return 1;
}

/* The reference's loop over points, their influence blocks (finest leaves
 * with data that overlap with radius 1, client.hpp:1264-1320) and the block
 * nodes within ddf_radius+1; mode 0 spreads (operators.hpp:996-1015), mode 1
 * interpolates (operators.hpp:844-861). */
static void ib_transfer(const oracle_tree* t, const oracle_ib* ib, int mode, const double* fin, double** fields,
    double* fout)
{
    const int    L = t->depth - 1, n = t->n, e = n - 2;
    const size_t n3 = (size_t)n * n * n;
    int          i, s, a, b, c, comp, d;
    if (mode == 1) memset(fout, 0, 3 * (size_t)ib->npoints * sizeof(double));
    for (i = 0; i < ib->npoints; ++i)
        for (s = level_first(t, L); s < t->nb && t->level[s] == L; ++s)
        {
            double p, X[3];
            if (!(t->flags[s] & 1u) || !ib_overlap(t, ib, i, s, 1)) continue;
            p = pow(2, t->level[s] - t->base_level);
            for (d = 0; d < 3; ++d) X[d] = ib->x[3 * i + d] * p / t->dx_base;
            for (c = 0; c < e; ++c)
                for (b = 0; b < e; ++b)
                    for (a = 0; a < e; ++a)
                    {
                        const int nc[3] = {t->bbase[3 * s] + a, t->bbase[3 * s + 1] + b, t->bbase[3 * s + 2] + c};
                        double    dist[3];
                        int       inside = 1;
                        size_t    q = s * n3 + (((size_t)c + 1) * n + (b + 1)) * n + (a + 1);
                        for (d = 0; d < 3; ++d)
                        {
                            dist[d] = nc[d] - X[d];
                            if (fabs(dist[d]) >= ib->ddf_radius + 1.0) inside = 0;
                        }
                        if (!inside) continue;
                        for (comp = 0; comp < 3; ++comp)
                        {
                            double w, v[3];
                            for (d = 0; d < 3; ++d) v[d] = dist[d] + (d == comp ? 0.0 : 0.5);
                            w = ddf1(ib->delta, v[0]) * ddf1(ib->delta, v[1]) * ddf1(ib->delta, v[2]);
                            if (mode == 0) fields[comp][q] += fin[3 * i + comp] * w * 1.0;
                            else
                                fout[3 * i + comp] += fields[comp][q] * w;
                        }
                    }
        }
}

void oracle_ib_smear(const oracle_tree* t, const oracle_ib* ib, const double* force, double* f0, double* f1, double* f2)
{
    double*      fl[3] = {f0, f1, f2};
    const size_t sz = (size_t)t->nb * t->n * t->n * t->n;
    int          c;
    for (c = 0; c < 3; ++c) memset(fl[c], 0, sz * sizeof(double));
    ib_transfer(t, ib, 0, force, fl, NULL);
}

void oracle_ib_project(const oracle_tree* t, const oracle_ib* ib, const double* f0, const double* f1,
    const double* f2, double* force)
{
    double* fl[3] = {(double*)f0, (double*)f1, (double*)f2};
    ib_transfer(t, ib, 1, NULL, fl, force);
}

/* points strictly inside a finest leaf block (ib rank != -1) */
void oracle_ib_owned(const oracle_tree* t, const oracle_ib* ib, int* owned)
{
    const int L = t->depth - 1;
    int       i, s;
    for (i = 0; i < ib->npoints; ++i)
    {
        owned[i] = 0;
        for (s = level_first(t, L); s < t->nb && t->level[s] == L; ++s)
            if ((t->flags[s] & 1u) && ib_overlap(t, ib, i, s, 1) && ib_overlap(t, ib, i, s, 0))
            {
                owned[i] = 1;
                break;
            }
    }
}

/* LinSysSolver::ET_H_S_E (linsys.hpp:259-319) */
int oracle_ib_apply(const oracle_tree* t, const oracle_ib* ib, const double* g, int nmax, const double* fin,
    double* fout, double alpha)
{
    const size_t sz = (size_t)t->nb * t->n * t->n * t->n;
    const int    L = t->depth - 1;
    double *     F[3], *A[3], *cell = (double*)calloc(sz, sizeof(double));
    size_t       q;
    int          c;
    for (c = 0; c < 3; ++c)
    {
        F[c] = (double*)calloc(sz, sizeof(double));
        A[c] = (double*)calloc(sz, sizeof(double));
    }
    oracle_ib_smear(t, ib, fin, F[0], F[1], F[2]);
    if (fabs(alpha) > 1e-14)
        for (c = 0; c < 3; ++c) oracle_apply_if_masked(t, F[c], F[c], alpha, 2, 0, 5);
    for (c = 0; c < 3; ++c) oracle_buffer_exchange(t, F[c], L);
    oracle_block_op(t, 2, F[0], F[1], F[2], cell, NULL, NULL, L, 1.0);
    oracle_clean_boundary(t, cell, L, 0, 1, 2);
    oracle_clean_boundary(t, cell, L, 1, 0, 2);
    oracle_apply_lgf_masked(t, g, nmax, cell, cell, 3, 0, 5);
    oracle_block_op(t, 1, cell, NULL, NULL, A[0], A[1], A[2], L, 1.0);
    for (c = 0; c < 3; ++c) oracle_buffer_exchange(t, A[c], L);
    for (c = 0; c < 3; ++c) oracle_clean_boundary(t, A[c], L, 1, 0, 2);
    for (c = 0; c < 3; ++c)
        for (q = 0; q < sz; ++q) F[c][q] += A[c][q] * -1.0;
    oracle_ib_project(t, ib, F[0], F[1], F[2], fout);
    for (c = 0; c < 3; ++c)
    {
        free(F[c]);
        free(A[c]);
    }
    free(cell);
    return 0;
}

/* LinSysSolver::CG_solve / dot / add (linsys.hpp:170-230, 500-531) */
int oracle_ib_cg(const oracle_tree* t, const oracle_ib* ib, const double* g, int nmax, const double* uc, double* f,
    double alpha, double tol, int maxit, double* residual)
{
    const int N = ib->npoints;
    double *  Ax = (double*)calloc(3 * N, sizeof(double)), *r = (double*)calloc(3 * N, sizeof(double));
    double *  p = (double*)calloc(3 * N, sizeof(double)), *Ap = (double*)calloc(3 * N, sizeof(double));
    int*      own = (int*)calloc(N, sizeof(int));
    int       i, d, k, iters = 0;
    double    rsold = 0, res = 0;
    oracle_ib_owned(t, ib, own);
    oracle_ib_apply(t, ib, g, nmax, f, Ax, alpha);
    for (i = 0; i < N; ++i)
        for (d = 0; d < 3; ++d) r[3 * i + d] = own[i] ? uc[3 * i + d] - Ax[3 * i + d] : 0.0;
    memcpy(p, r, 3 * N * sizeof(double));
    for (i = 0; i < N; ++i)
        if (own[i])
            for (d = 0; d < 3; ++d) rsold += r[3 * i + d] * r[3 * i + d];
    for (k = 0; k < maxit; ++k)
    {
        double pAp = 0, a, rsnew = 0, f2 = 0;
        oracle_ib_apply(t, ib, g, nmax, p, Ap, alpha);
        for (i = 0; i < N; ++i)
            if (own[i])
                for (d = 0; d < 3; ++d) pAp += p[3 * i + d] * Ap[3 * i + d];
        if (pAp == 0.0) break;
        a = rsold / pAp;
        for (i = 0; i < N; ++i)
            if (own[i])
                for (d = 0; d < 3; ++d)
                {
                    f[3 * i + d] = f[3 * i + d] * 1.0 + p[3 * i + d] * a;
                    r[3 * i + d] = r[3 * i + d] * 1.0 + Ap[3 * i + d] * -a;
                }
        for (i = 0; i < N; ++i)
            if (own[i])
                for (d = 0; d < 3; ++d)
                {
                    rsnew += r[3 * i + d] * r[3 * i + d];
                    f2 += f[3 * i + d] * f[3 * i + d];
                }
        iters = k + 1;
        res = sqrt(rsnew / f2);
        if (res < tol) break;
        for (i = 0; i < N; ++i)
            if (own[i])
                for (d = 0; d < 3; ++d) p[3 * i + d] = p[3 * i + d] * (rsnew / rsold) + r[3 * i + d] * 1.0;
        rsold = rsnew;
    }
    if (residual) *residual = res;
    free(Ax);
    free(r);
    free(p);
    free(Ap);
    free(own);
    return iters;
}

/* ---- IF-HERK time step (ifherk.hpp:1078-1200) ------------------------------
 * Without the immersed boundary (no IB points: lin_sys_with_ib_solve,
 * 1451-1517, reduces to the projection of lin_sys_solve, 1425-1449, with the
 * IF threshold 1e-14).  Fields are [nb][n^3]; face and edge vectors are three
 * arrays (component c = field index c). */

/* Ifherk::clean (1306-1350): every block zero, or with non_leaf_only the
 * one-point ring of the leaves and every non-leaf block */
static void hk_clean(const oracle_tree* t, double* f, int non_leaf_only)
{
    const int    n = t->n;
    const size_t n3 = (size_t)n * n * n;
    int          s, i, j, k;
    for (s = 0; s < t->nb; ++s)
    {
        double* b = f + s * n3;
        if (non_leaf_only && (t->flags[s] & 1u))
        {
            for (k = 0; k < n; ++k)
                for (j = 0; j < n; ++j)
                    for (i = 0; i < n; ++i)
                        if (i == 0 || j == 0 || k == 0 || i == n - 1 || j == n - 1 || k == n - 1)
                            b[((size_t)k * n + j) * n + i] = 0.0;
        }
        else
            memset(b, 0, n3 * sizeof(double));
    }
}

/* Ifherk::copy (1944-1960): to = from * scale on every node */
static void hk_copy(const oracle_tree* t, const double* from, double* to, double scale)
{
    const size_t len = (size_t)t->nb * t->n * t->n * t->n;
    size_t       q;
    for (q = 0; q < len; ++q) to[q] = from[q] * scale;
}

/* Ifherk::add (1877-1893): to += from * scale */
static void hk_add(const oracle_tree* t, const double* from, double* to, double scale)
{
    const size_t len = (size_t)t->nb * t->n * t->n * t->n;
    size_t       q;
    for (q = 0; q < len; ++q) to[q] = to[q] + from[q] * scale;
}

/* smooth2zero's ramp (operators.hpp:668-675) */
static double hk_ramp(double x)
{
    const double fac = 10.0, shift = 0.2;
    const double c = 1 - (0.5 + 0.5 * tanh(fac * (1 - shift)));
    return ((0.5 + 0.5 * tanh(fac * (x - shift))) + c);
}

/* Operator::smooth2zero (operators.hpp:653-760): every node of the block is
 * multiplied by the largest ramp value over the axes on which neighbour
 * ngb (x fastest in 3x3x3) lies off the block; the node's position on an
 * axis is its storage index / (n - 1). */
static void hk_smooth2zero(double* blk, int n, int ngb)
{
    const int x = ngb % 3, y = (ngb / 3) % 3, z = (ngb / 9) % 3;
    int       i, j, k;
    for (k = 0; k < n; ++k)
        for (j = 0; j < n; ++j)
            for (i = 0; i < n; ++i)
            {
                const double px = (double)i / (double)(n - 1), py = (double)j / (double)(n - 1),
                             pz = (double)k / (double)(n - 1);
                double square = 0.0, r;
                int    c = 0;
                if (z == 0) { r = hk_ramp(pz); square = square > r ? square : r; ++c; }
                else if (z == 2) { r = hk_ramp(1 - pz); square = square > r ? square : r; ++c; }
                if (y == 0) { r = hk_ramp(py); square = square > r ? square : r; ++c; }
                else if (y == 2) { r = hk_ramp(1 - py); square = square > r ? square : r; ++c; }
                if (x == 0) { r = hk_ramp(px); square = square > r ? square : r; ++c; }
                else if (x == 2) { r = hk_ramp(1 - px); square = square > r ? square : r; ++c; }
                if (c > 0) blk[((size_t)k * n + j) * n + i] = blk[((size_t)k * n + j) * n + i] * square;
            }
}

/* Ifherk::clean_leaf_correction_boundary (1352-1412), one rank: with
 * leaf_only the correction blocks of level l are zeroed; at the base level
 * every block is ramped towards each neighbour that is missing (or, with
 * leaf_only, a correction block), neighbours in ascending index. */
void oracle_smooth_boundary(const oracle_tree* t, double* f, int l, int leaf_only)
{
    const int    n = t->n;
    const size_t n3 = (size_t)n * n * n;
    int          s, i;
    if (leaf_only)
        for (s = level_first(t, l); s < t->nb && t->level[s] == l; ++s)
            if (t->flags[s] & 2u) memset(f + s * n3, 0, n3 * sizeof(double));
    if (l != t->base_level) return;
    for (s = level_first(t, l); s < t->nb && t->level[s] == l; ++s)
        for (i = 0; i < 27; ++i)
        {
            const int nb = t->neighbor[27 * s + i];
            if (nb < 0 || (leaf_only && (t->flags[nb] & 2u))) hk_smooth2zero(f + s * n3, n, i);
        }
}

/* Operator::nonlinear (operators.hpp:2455-2500): dst = 0.25 (omega x u) on
 * the interior nodes of every block at level l, face velocity u, edge
 * vorticity omega, face result; then every node of the block times scale
 * (Ifherk::nonlinear, ifherk.hpp:1676-1686). */
void oracle_nonlinear(const oracle_tree* t, const double* u0, const double* u1, const double* u2, const double* w0,
    const double* w1, const double* w2, double* d0, double* d1, double* d2, int l, double scale)
{
    const int    n = t->n;
    const size_t n3 = (size_t)n * n * n, X = 1, Y = n, Z = (size_t)n * n;
    int          b, i, j, k;
    size_t       q;
    for (b = level_first(t, l); b < t->nb && t->level[b] == l; ++b)
    {
        for (k = 1; k < n - 1; ++k)
            for (j = 1; j < n - 1; ++j)
                for (i = 1; i < n - 1; ++i)
                {
                    q = (size_t)b * n3 + (k * Z + j * Y + i * X);
                    d0[q] = 0.25 * (+w1[q] * (+u2[q] + u2[q - X]) + w1[q + Z] * (+u2[q + Z] + u2[q + Z - X]) -
                                       w2[q] * (+u1[q] + u1[q - X]) - w2[q + Y] * (+u1[q + Y] + u1[q + Y - X]));
                    d1[q] = 0.25 * (+w2[q] * (+u0[q] + u0[q - Y]) + w2[q + X] * (+u0[q + X] + u0[q + X - Y]) -
                                       w0[q] * (+u2[q] + u2[q - Y]) - w0[q + Z] * (+u2[q + Z] + u2[q + Z - Y]));
                    d2[q] = 0.25 * (+w0[q] * (+u1[q] + u1[q - Z]) + w0[q + Y] * (+u1[q + Y] + u1[q + Y - Z]) -
                                       w1[q] * (+u0[q] + u0[q - Z]) - w1[q + X] * (+u0[q + X] + u0[q + X - Z]));
                }
        for (q = (size_t)b * n3; q < (size_t)(b + 1) * n3; ++q)
        {
            d0[q] *= scale;
            d1[q] *= scale;
            d2[q] *= scale;
        }
    }
}

typedef struct hk_state
{
    const oracle_tree* t;
    const double*      g;
    int                nmax;
    double *q[3], *r[3], *gi[3], *w1[3], *w2[3], *ui[3], *fa[3], *ea[3], *sf[3], *di, *ca;
    /* immersed boundary (NULL: none) */
    const oracle_ib* ib;
    double *         fa2[3], *ca2, *force, *uc;
    double*          force_prev; /* [3 stages][3N] */
    const double*    bc;         /* [3 stages][3] */
    const double*    bc_pts;     /* [3 stages][N][3] or NULL */
    double           cg_tol;
    int              cg_maxit;
} hk_state;

/* Ifherk::up_and_down (991-997): clean<Field>(true), then per component the
 * face-mesh coarsening (up, 999-1004) and the interpolation into the
 * correction blocks (down_to_correction, 1006-1013). */
/* Ifherk::up (999-1004) after clean<Field>(true): every component coarsened */
static void hk_up_mesh(const oracle_tree* t, double* const f[3], int mesh)
{
    ostate  o_;
    ostate* o = &o_;
    int     c;
    ostate_init(o, t, 0, 5);
    o->mesh = mesh;
    for (c = 0; c < 3; ++c) hk_clean(t, f[c], 1);
    for (c = 0; c < 3; ++c)
    {
        o->comp = c;
        source_coarsify(o, f[c]);
    }
    ostate_free(o);
}

/* Ifherk::down_to_correction (1006-1013): per component, each level's halo
 * fill (of every component, as buffer_exchange<Field>) before its
 * interpolation into the correction children */
static void hk_down_mesh(const oracle_tree* t, double* const f[3], int mesh)
{
    ostate  o_;
    ostate* o = &o_;
    int     c, l, s;
    ostate_init(o, t, 0, 5);
    o->mesh = mesh;
    for (c = 0; c < 3; ++c)
    {
        o->comp = c;
        for (l = t->depth - 2; l >= t->base_level; --l)
        {
            int e;
            for (e = 0; e < 3; ++e) buffer_exchange(o, f[e], l);
            for (s = o->lvl_begin[l]; s < o->lvl_begin[l + 1]; ++s) cc_intrp_node(o, f[c], f[c], s, 1);
        }
    }
    ostate_free(o);
}

static void hk_up_and_down_mesh(const oracle_tree* t, double* const f[3], int mesh)
{
    hk_up_mesh(t, f, mesh);
    hk_down_mesh(t, f, mesh);
}

static void hk_up_and_down(const hk_state* h, double* const f[3]) { hk_up_and_down_mesh(h->t, f, 1); }

void oracle_up_and_down(const oracle_tree* t, double* f0, double* f1, double* f2, int mesh)
{
    double* const f[3] = {f0, f1, f2};
    hk_up_and_down_mesh(t, f, mesh);
}

/* Ifherk::nonlinear (1637-1703) with a frame velocity that is constant over
 * the domain (add_field_expression, operators.hpp:2985-3002) */
static void hk_nonlinear(const hk_state* h, double* const src[3], double* const dst[3], double scale,
    const double frame[3])
{
    const oracle_tree* t = h->t;
    int                c, l;
    const size_t       len = (size_t)t->nb * t->n * t->n * t->n;
    for (c = 0; c < 3; ++c)
    {
        hk_clean(t, h->ea[c], 0);
        hk_clean(t, dst[c], 0);
        hk_clean(t, h->fa[c], 0);
    }
    for (l = t->base_level; l < t->depth; ++l)
    {
        for (c = 0; c < 3; ++c) oracle_buffer_exchange(t, src[c], l);
        oracle_block_op(t, 3, src[0], src[1], src[2], h->ea[0], h->ea[1], h->ea[2], l, 1.0);
    }
    for (c = 0; c < 3; ++c)
    {
        size_t q;
        hk_copy(t, src[c], h->fa[c], 1.0);
        for (q = 0; q < len; ++q) h->fa[c][q] += frame[c] * -1.0;
    }
    for (l = t->base_level; l < t->depth; ++l)
    {
        for (c = 0; c < 3; ++c) oracle_buffer_exchange(t, h->ea[c], l);
        for (c = 0; c < 3; ++c) oracle_smooth_boundary(t, h->ea[c], l, 0);
        oracle_nonlinear(t, h->fa[0], h->fa[1], h->fa[2], h->ea[0], h->ea[1], h->ea[2], dst[0], dst[1], dst[2], l,
            scale);
    }
}

/* Ifherk::divergence (1738-1764) */
static void hk_divergence(const hk_state* h, double* const src[3], double* dst)
{
    const oracle_tree* t = h->t;
    int                c, l;
    hk_up_and_down(h, src);
    for (l = t->base_level; l < t->depth; ++l)
    {
        for (c = 0; c < 3; ++c) oracle_buffer_exchange(t, src[c], l);
        oracle_block_op(t, 2, src[0], src[1], src[2], dst, NULL, NULL, l, 1.0);
        oracle_smooth_boundary(t, dst, l, 1);
    }
}

/* Ifherk::gradient (1766-1793), scale 1 */
static void hk_gradient(const hk_state* h, const double* src, double* const dst[3])
{
    const oracle_tree* t = h->t;
    int                c, l;
    for (c = 0; c < 3; ++c) hk_clean(t, dst[c], 0);
    for (l = t->base_level; l < t->depth; ++l)
    {
        oracle_block_op(t, 1, src, NULL, NULL, dst[0], dst[1], dst[2], l, 1.0);
        for (c = 0; c < 3; ++c) oracle_buffer_exchange(t, dst[c], l);
    }
}

static void hk_if(const hk_state* h, double* const src[3], double* const dst[3], double alpha)
{
    int c;
    for (c = 0; c < 3; ++c) if_solve(h->t, src[c], dst[c], alpha, 0, 5, 1, c);
}

/* lin_sys_with_ib_solve (1451-1517) with no immersed-boundary points: the
 * IB2xIB solve, the force solve and the pressure correction act on zero
 * forces and leave d_i, face_aux and r_i as they are. */
/* Operator::levelDivergence (operators.hpp:522-538) at level l */
static void hk_level_divergence(const oracle_tree* t, double* const src[3], double* dst, int l)
{
    int c;
    for (c = 0; c < 3; ++c) oracle_buffer_exchange(t, src[c], l);
    oracle_block_op(t, 2, src[0], src[1], src[2], dst, NULL, NULL, l, 1.0);
    oracle_clean_boundary(t, dst, l, 0, 1, 2);
}

/* smearing (linsys.hpp:350-376) of the IB force, adding into the fields */
static void hk_smear_add(const hk_state* h, double* const f[3])
{
    double* fl[3] = {f[0], f[1], f[2]};
    ib_transfer(h->t, h->ib, 0, h->force, fl, NULL);
}

/* The immersed-boundary part of lin_sys_with_ib_solve (ifherk.hpp:1465-1494):
 * face_aux2 = r_i - grad d_i through the IB2xIB integrating factor, the
 * force solve (LinSysSolver::ib_solve, linsys.hpp:113-129: the velocity at
 * the points less the boundary velocity, CG_solve from the stage's previous
 * force), then pressure_correction (131-145). */
static void hk_ib_solve(const hk_state* h, double alpha, int stage, int write_prev)
{
    const oracle_tree* t = h->t;
    const int          N = h->ib->npoints, L = t->depth - 1;
    int                c, i;
    for (c = 0; c < 3; ++c)
    {
        hk_copy(t, h->r[c], h->fa2[c], 1.0);
        hk_add(t, h->fa[c], h->fa2[c], -1.0);
    }
    if (fabs(alpha) > 1e-14)
        for (c = 0; c < 3; ++c) oracle_apply_if_masked(t, h->fa2[c], h->fa2[c], alpha, 2, 0, 5);
    memcpy(h->force, h->force_prev + (size_t)3 * N * (stage - 1), 3 * (size_t)N * sizeof(double));
    oracle_ib_project(t, h->ib, h->fa2[0], h->fa2[1], h->fa2[2], h->uc);
    for (i = 0; i < N; ++i)
        for (c = 0; c < 3; ++c)
            h->uc[3 * i + c] -= h->bc_pts ? h->bc_pts[(size_t)3 * N * (stage - 1) + 3 * i + c]
                                          : h->bc[3 * (stage - 1) + c];
    for (c = 0; c < 3; ++c) hk_clean(t, h->fa2[c], 0);
    oracle_ib_cg(t, h->ib, h->g, h->nmax, h->uc, h->force, alpha, h->cg_tol, h->cg_maxit, NULL);
    if (write_prev) memcpy(h->force_prev + (size_t)3 * N * (stage - 1), h->force, 3 * (size_t)N * sizeof(double));
    /* pressure_correction<d_i> */
    for (c = 0; c < 3; ++c) hk_clean(t, h->fa2[c], 0);
    hk_clean(t, h->ca2, 0);
    hk_smear_add(h, h->fa2);
    hk_level_divergence(t, h->fa2, h->ca2, L);
    oracle_apply_lgf_masked(t, h->g, h->nmax, h->ca2, h->ca2, 4, 0, 5);
    hk_add(t, h->ca2, h->di, -1.0);
}

static void hk_lin_sys(const hk_state* h, double alpha, int stage, int write_prev)
{
    const oracle_tree* t = h->t;
    int                c;
    hk_divergence(h, h->r, h->ca);
    {
        ostate  o_;
        ostate* o = &o_;
        ostate_init(o, t, 0, 5);
        o->ker.is_if = 0;
        o->ker.g = h->g;
        o->ker.nmax = h->nmax;
        lgf_solve(o, h->ca, h->di);
        ostate_free(o);
    }
    hk_gradient(h, h->di, h->fa);
    if (h->ib)
    {
        hk_ib_solve(h, alpha, stage, write_prev);
        hk_gradient(h, h->di, h->fa);
        hk_smear_add(h, h->fa);
    }
    for (c = 0; c < 3; ++c) hk_add(t, h->fa[c], h->r[c], -1.0);
    if (fabs(alpha) > 1e-14) hk_if(h, h->r, h->ui, alpha);
    else
        for (c = 0; c < 3; ++c) hk_copy(t, h->r[c], h->ui[c], 1.0);
}

/* Ifherk::pad_velocity<u, u>(true) (1521-1575): the velocity coarsened, its
 * curl on every non-correction block of every level, ramped at the base
 * level, the base-level stream function by the STREAM solve of each
 * component, the velocity on the base level from its curl transpose (scale
 * -1), then interpolated into the correction blocks. */
static void hk_pad_velocity(const hk_state* h, double* const u[3])
{
    const oracle_tree* t = h->t;
    const int          n = t->n;
    const size_t       n3 = (size_t)n * n * n, len = (size_t)t->nb * n3;
    double*            tmp[3];
    int                c, l, s;
    hk_up_mesh(t, u, 1);
    for (c = 0; c < 3; ++c)
    {
        hk_clean(t, h->ea[c], 0);
        hk_clean(t, h->sf[c], 0);
        tmp[c] = (double*)calloc(len, sizeof(double));
    }
    for (l = t->base_level; l < t->depth; ++l)
    {
        for (c = 0; c < 3; ++c) oracle_buffer_exchange(t, u[c], l);
        oracle_block_op(t, 3, u[0], u[1], u[2], tmp[0], tmp[1], tmp[2], l, 1.0);
        for (s = level_first(t, l); s < t->nb && t->level[s] == l; ++s)
            if (!(t->flags[s] & 2u))
                for (c = 0; c < 3; ++c) memcpy(h->ea[c] + s * n3, tmp[c] + s * n3, n3 * sizeof(double));
    }
    for (c = 0; c < 3; ++c)
    {
        free(tmp[c]);
        oracle_smooth_boundary(t, h->ea[c], t->base_level, 1);
    }
    for (c = 0; c < 3; ++c) oracle_apply_lgf_masked(t, h->g, h->nmax, h->ea[c], h->sf[c], 0, 0, 5);
    oracle_block_op(t, 4, h->sf[0], h->sf[1], h->sf[2], u[0], u[1], u[2], t->base_level, -1.0);
    hk_down_mesh(t, u, 1);
}

/* Ifherk::adapt_corr_time_step (1206-1304): the three stages with alpha 0,
 * no nonlinear term and the stage integrating factor of alpha 0 */
static void hk_adapt_corr(const hk_state* h, double* const u[3])
{
    const oracle_tree* t = h->t;
    int                c;
    hk_up_and_down(h, u);
    for (c = 0; c < 3; ++c) hk_copy(t, u[c], h->q[c], 1.0);
    for (c = 0; c < 3; ++c)
    {
        hk_clean(t, h->gi[c], 0);
        hk_clean(t, h->fa[c], 0);
    }
    hk_clean(t, h->di, 0);
    hk_clean(t, h->ca, 0);
    for (c = 0; c < 3; ++c) hk_copy(t, h->q[c], h->r[c], 1.0);
    hk_lin_sys(h, 0.0, 1, 0);
    for (c = 0; c < 3; ++c) hk_clean(t, h->r[c], 0);
    hk_clean(t, h->di, 0);
    hk_clean(t, h->ca, 0);
    for (c = 0; c < 3; ++c) hk_add(t, h->q[c], h->r[c], 1.0);
    hk_up_and_down(h, h->ui);
    hk_lin_sys(h, 0.0, 2, 0);
    hk_clean(t, h->di, 0);
    hk_clean(t, h->ca, 0);
    for (c = 0; c < 3; ++c)
    {
        hk_clean(t, h->w2[c], 0);
        hk_copy(t, h->q[c], h->r[c], 1.0);
    }
    hk_if(h, h->r, h->r, 0.0);
    hk_up_and_down(h, h->ui);
    hk_lin_sys(h, 0.0, 3, 0);
    for (c = 0; c < 3; ++c) hk_copy(t, h->ui[c], u[c], 1.0);
}

static int hk_step(const oracle_tree* t, const oracle_ib* ib, const double* g, int nmax, double* u0, double* u1,
    double* u2, double* p, double dt, double Re, const double* frame, int refresh, const double* bc,
    const double* bc_pts, double* force_prev, double cg_tol, int cg_maxit);

int oracle_ifherk_step(const oracle_tree* t, const double* g, int nmax, double* u0, double* u1, double* u2,
    double* p, double dt, double Re, const double* frame, int refresh)
{
    return hk_step(t, NULL, g, nmax, u0, u1, u2, p, dt, Re, frame, refresh, NULL, NULL, NULL, 0.0, 0);
}

int oracle_ifherk_step_ib(const oracle_tree* t, const oracle_ib* ib, const double* g, int nmax, double* u0,
    double* u1, double* u2, double* p, double dt, double Re, const double* frame, int refresh, const double* bc,
    double* force_prev, double cg_tol, int cg_maxit)
{
    if (!ib || !force_prev) return 1;
    return hk_step(t, ib, g, nmax, u0, u1, u2, p, dt, Re, frame, refresh, bc, NULL, force_prev, cg_tol, cg_maxit);
}

int oracle_ifherk_step_ib_points(const oracle_tree* t, const oracle_ib* ib, const double* g, int nmax, double* u0,
    double* u1, double* u2, double* p, double dt, double Re, const double* frame, int refresh,
    const double* bc_points, double* force_prev, double cg_tol, int cg_maxit)
{
    if (!ib || !force_prev || !bc_points) return 1;
    return hk_step(t, ib, g, nmax, u0, u1, u2, p, dt, Re, frame, refresh, NULL, bc_points, force_prev, cg_tol,
        cg_maxit);
}

static int hk_step(const oracle_tree* t, const oracle_ib* ib, const double* g, int nmax, double* u0, double* u1,
    double* u2, double* p, double dt, double Re, const double* frame, int refresh, const double* bc,
    const double* bc_pts, double* force_prev, double cg_tol, int cg_maxit)
{
    static const double a_[6] = {1.0 / 3, -1.0, 2.0, 0.0, 0.75, 0.25};
    static const double c_[4] = {0.0, 1.0 / 3, 1.0, 1.0};
#define COEFF_A(i, j) a_[(i) * ((i)-1) / 2 + (j)-1]
    const size_t len = (size_t)t->nb * t->n * t->n * t->n;
    double* const u[3] = {u0, u1, u2};
    double        zero[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, alpha[3], tmp;
    hk_state      h;
    int           c;
    double*       pool;
    if (!frame) frame = zero;
    memset(&h, 0, sizeof(h));
    h.t = t;
    h.g = g;
    h.nmax = nmax;
    pool = (double*)calloc(len * 33, sizeof(double));
    if (!pool) return 1;
    if (ib)
    {
        h.ib = ib;
        for (c = 0; c < 3; ++c) h.fa2[c] = pool + len * (29 + c);
        h.ca2 = pool + len * 32;
        h.force = (double*)calloc(3 * (size_t)ib->npoints, sizeof(double));
        h.uc = (double*)calloc(3 * (size_t)ib->npoints, sizeof(double));
        h.force_prev = force_prev;
        h.bc = bc ? bc : zero;
        h.bc_pts = bc_pts;
        h.cg_tol = cg_tol;
        h.cg_maxit = cg_maxit;
    }
    for (c = 0; c < 3; ++c)
    {
        h.q[c] = pool + len * (0 + c);
        h.r[c] = pool + len * (3 + c);
        h.gi[c] = pool + len * (6 + c);
        h.w1[c] = pool + len * (9 + c);
        h.w2[c] = pool + len * (12 + c);
        h.ui[c] = pool + len * (15 + c);
        h.fa[c] = pool + len * (18 + c);
        h.ea[c] = pool + len * (21 + c);
        h.sf[c] = pool + len * (26 + c);
    }
    h.di = pool + len * 24;
    h.ca = pool + len * 25;
    /* update_marching_parameters (ifherk.hpp:164-175) */
    tmp = Re * t->dx_base * t->dx_base / dt;
    alpha[0] = (c_[1] - c_[0]) / tmp;
    alpha[1] = (c_[2] - c_[1]) / tmp;
    alpha[2] = (c_[3] - c_[2]) / tmp;

    /* time_step (1093-1116): refresh the velocity from the stream function
     * (then, with use_adaptation_correction, the alpha-0 correction step),
     * or up_and_down */
    if (refresh)
    {
        hk_pad_velocity(&h, u);
        if (refresh == 1) hk_adapt_corr(&h, u);
    }
    else
        hk_up_and_down(&h, u);
    for (c = 0; c < 3; ++c) hk_copy(t, u[c], h.q[c], 1.0);

    /* stage 1 */
    for (c = 0; c < 3; ++c)
    {
        hk_clean(t, h.gi[c], 0);
        hk_clean(t, h.fa[c], 0);
    }
    hk_clean(t, h.di, 0);
    hk_clean(t, h.ca, 0);
    hk_nonlinear(&h, u, h.gi, COEFF_A(1, 1) * (-dt), frame);
    for (c = 0; c < 3; ++c)
    {
        hk_copy(t, h.q[c], h.r[c], 1.0);
        hk_add(t, h.gi[c], h.r[c], 1.0);
    }
    hk_lin_sys(&h, alpha[0], 1, 1);

    /* stage 2 */
    for (c = 0; c < 3; ++c) hk_clean(t, h.r[c], 0);
    hk_clean(t, h.di, 0);
    hk_clean(t, h.ca, 0);
    for (c = 0; c < 3; ++c)
    {
        hk_add(t, h.gi[c], h.fa[c], -1.0);
        hk_copy(t, h.fa[c], h.w1[c], -1.0 / dt / COEFF_A(1, 1));
    }
    hk_if(&h, h.q, h.q, alpha[0]);
    hk_if(&h, h.w1, h.w1, alpha[0]);
    for (c = 0; c < 3; ++c)
    {
        hk_add(t, h.q[c], h.r[c], 1.0);
        hk_add(t, h.w1[c], h.r[c], dt * COEFF_A(2, 1));
    }
    hk_up_and_down(&h, h.ui);
    hk_nonlinear(&h, h.ui, h.gi, COEFF_A(2, 2) * (-dt), frame + 3);
    for (c = 0; c < 3; ++c) hk_add(t, h.gi[c], h.r[c], 1.0);
    hk_lin_sys(&h, alpha[1], 2, 1);

    /* stage 3 */
    hk_clean(t, h.di, 0);
    hk_clean(t, h.ca, 0);
    for (c = 0; c < 3; ++c) hk_clean(t, h.w2[c], 0);
    for (c = 0; c < 3; ++c)
    {
        hk_add(t, h.gi[c], h.fa[c], -1.0);
        hk_copy(t, h.fa[c], h.w2[c], -1.0 / dt / COEFF_A(2, 2));
        hk_copy(t, h.q[c], h.r[c], 1.0);
        hk_add(t, h.w1[c], h.r[c], dt * COEFF_A(3, 1));
        hk_add(t, h.w2[c], h.r[c], dt * COEFF_A(3, 2));
    }
    hk_if(&h, h.r, h.r, alpha[1]);
    hk_up_and_down(&h, h.ui);
    hk_nonlinear(&h, h.ui, h.gi, COEFF_A(3, 3) * (-dt), frame + 6);
    for (c = 0; c < 3; ++c) hk_add(t, h.gi[c], h.r[c], 1.0);
    hk_lin_sys(&h, alpha[2], 3, 1);

    for (c = 0; c < 3; ++c) hk_copy(t, h.ui[c], u[c], 1.0);
    if (p) hk_copy(t, h.di, p, 1.0 / COEFF_A(3, 3) / dt);
#undef COEFF_A
    free(pool);
    free(h.force);
    free(h.uc);
    return 0;
}

/* ---- Screened 2-D LGF of one spanwise Fourier mode --------------------------
 * Helmholtz::get for Dim = 2 (include/iblgf/lgf/helmholtz.hpp:143-153): the
 * trapezoid integral of (1 - cos(m x) K^-n) / (K - 1/K) / (2 pi) over a period
 * (helmholtz_lookup.hpp:60-68, 170-181) minus origin_val
 * (helmholtz_lookup.hpp:114-118), n >= m the larger offset.  The first term's
 * integral is origin_val itself, so the value is -G(n, m) with
 *   G(n, m) = 1/pi int_0^pi cos(m x) K^-n / sqrt(a^2 - 4) dx,
 * a = 4 + c^2 - 2 cos x, K = (a + sqrt(a^2 - 4)) / 2, evaluated here by the
 * trapezoid rule with `nodes` intervals on the half period (the integrand is
 * even).  Pinned in tests against origin_val's hypergeometric series, an
 * adaptive quadrature of the same integral and the lattice identity
 * (Lap_h - c^2) value = delta. */
double oracle_helm2d_lgf(double c, int i, int j, int nodes)
{
    int n = abs(i), m = abs(j);
    if (n < m)
    {
        int t = n;
        n = m;
        m = t;
    }
    const double h = O_PI / nodes;
    double acc = 0.0;
    for (int p = 0; p <= nodes; ++p)
    {
        const double x = p * h;
        const double sh = sin(0.5 * x);
        const double e = c * c + 4.0 * sh * sh;
        const double root = sqrt(e * (e + 4.0));
        double f = cos(m * x) * exp(-n * log1p(0.5 * (e + root))) / root;
        if (p == 0 || p == nodes) f *= 0.5;
        acc += f;
    }
    return -acc / nodes;
}

/* Convolution<2> of one mode over a uniform nx x ny patch
 * (include/iblgf/utilities/convolution.hpp:259-260 summed over every block
 * pair of one level): dst[t] = scale * sum_s table[|tj-sj|][|ti-si|] src[s],
 * sources in index order. */
void oracle_helm2d_convolve(const double* table, int nx, int ny, const double* src, double* dst, double scale)
{
    const int per = nx * ny;
    for (int t = 0; t < per; ++t)
    {
        const int ti = t % nx, tj = t / nx;
        double acc = 0.0;
        for (int s = 0; s < per; ++s)
        {
            const int si = s % nx, sj = s / nx;
            acc += table[abs(tj - sj) * nx + abs(ti - si)] * src[s];
        }
        dst[t] = scale * acc;
    }
}
