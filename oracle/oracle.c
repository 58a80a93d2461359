/*
 * oracle.c -- CPU restatement of spaghetty's tensor-sort path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the spaghetty_amd product links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker (or as the
 * timed CPU baseline), never as the thing measured or shipped.
 *
 * Contents (each function cites the reference file:line it restates):
 *   oracle_old_sort_4 / oracle_old_sortacc_4   trunk/old_sort.f:33-95, 512-574
 *   oracle_old_sort_6 / oracle_old_sortacc_6   trunk/old_sort.f:97-234, 576-713
 *   oracle_sort_n                              ranks 0/2/8 by the same rule,
 *                                              trunk/old_sort.f:1-31, 236-478
 *   oracle_trans                               the generated library kernel
 *       trans_<P>_loop_<L>_<omp|ser>_any_f     trunk/generator.py:107-159 (4D),
 *       and the five-case (f,g) form           trunk/generator2.py:133-443,
 *                                              trunk/generator6.py:137-220
 *   oracle_init_kat                            trunk/tester_futil.F:139-185
 *   oracle_fill_uniform                        seeded uniform(-1,1) inputs
 *                                              (splitmix64; defined here, not
 *                                              in the reference)
 *
 * Pinning (see tests/test_oracle.py):
 *   - every old_sort restatement is compared bit-for-bit with the reference's
 *     own old_sort.f compiled from /root/reference (oracle/_ref), on the
 *     reference testers' integer KAT inputs and on seeded uniform inputs, for
 *     all 24 4D and all 720 6D permutations, factors {1,-1,0.5,37}, sort and
 *     sort-accumulate;
 *   - against committed golden digests made from oracle/_ref
 *     (tests/golden/make_golden.py), so the check still runs where the
 *     reference is absent;
 *   - against the NumPy closed form f * U.reshape(dims).transpose(perm-1),
 *     inverse-permutation round trips and brute force on tiny inputs.
 *
 * Build: -O3 -ffp-contract=off (the accumulate rounds twice: y + (x*s)),
 * -fopenmp for the timed CPU baseline of the generated OpenMP kernels.
 */
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* FP contraction is disabled by the build flags (-ffp-contract=off). */

/* ------------------------------------------------------------------------ */
/* NWChem index rule.  Dims jd(1..R) slowest->fastest; input element with   */
/* 1-based coordinates id(1..R) is at ia = id(R)+jd(R)*(id(R-1)-1+...)      */
/* (old_sort.f:51); the output is at                                        */
/* ib = id(p_R)+jd(p_R)*(id(p_{R-1})-1+jd(p_{R-1})*(...)) (old_sort.f:52).   */
/* ------------------------------------------------------------------------ */

#define ORACLE_MAX_RANK 8

/* Output strides (0-based elements) of every input axis for permutation
 * perm (1-based, NWChem order): the output's fastest axis is input axis
 * perm[R-1]; old_sort.f:52. */
static void out_strides(int R, const int *dims, const int *perm, int64_t *ostr)
{
    int64_t s = 1;
    for (int q = R - 1; q >= 0; --q) {
        ostr[perm[q] - 1] = s;
        s *= dims[perm[q] - 1];
    }
}

static void in_strides(int R, const int *dims, int64_t *istr)
{
    int64_t s = 1;
    for (int a = R - 1; a >= 0; --a) {
        istr[a] = s;
        s *= dims[a];
    }
}

/* Walk every element of the input in the loop order `order` (1-based axis
 * numbers outermost -> innermost) and apply y(ib) = x(ia)*s   (mode 0)
 *                                        y(ib) = y(ib) + x(ia)*s   (mode 1)
 *                                        y(ib) = g*y(ib) + f*x(ia) (mode 2).
 * The loop order only changes the order of writes, never a value: every
 * output element is written exactly once. */
static void walk(const double *x, double *y, int R, const int *dims,
                 const int *perm, const int *order, double f, double g, int mode)
{
    int64_t istr[ORACLE_MAX_RANK], ostr[ORACLE_MAX_RANK];
    int64_t n = 1;
    for (int a = 0; a < R; ++a) n *= dims[a];
    if (n <= 0) return;
    in_strides(R, dims, istr);
    out_strides(R, dims, perm, ostr);
    int ax[ORACLE_MAX_RANK];
    for (int t = 0; t < R; ++t) ax[t] = order[t] - 1;
    int idx[ORACLE_MAX_RANK] = {0};
    int64_t ia = 0, ib = 0;
    const int inner = ax[R - 1];
    const int64_t is = istr[inner], os = ostr[inner];
    const int nin = dims[inner];
    for (;;) {
        const double *xp = x + ia;
        double *yp = y + ib;
        if (mode == 0) {
            for (int t = 0; t < nin; ++t) yp[t * os] = xp[t * is] * f;
        } else if (mode == 1) {
            for (int t = 0; t < nin; ++t) {
                double p = xp[t * is] * f;
                yp[t * os] = yp[t * os] + p;
            }
        } else {
            for (int t = 0; t < nin; ++t) {
                double p = xp[t * is] * f;
                double q = yp[t * os] * g;
                yp[t * os] = q + p;
            }
        }
        /* odometer over the outer loops, innermost of them first */
        int lev = R - 2;
        for (; lev >= 0; --lev) {
            int a = ax[lev];
            idx[a]++;
            ia += istr[a];
            ib += ostr[a];
            if (idx[a] < dims[a]) break;
            ia -= istr[a] * dims[a];
            ib -= ostr[a] * dims[a];
            idx[a] = 0;
        }
        if (lev < 0) break;
    }
}

/* old_sort_4 chooses its loop order from l (old_sort.f:45,61,77):
 * l in {3,4}: j1 j2 j3 j4; l = 2: j1 j3 j2 j4; else: j2 j3 j1 j4. */
static void old4_order(int l, int *order)
{
    static const int o34[4] = {1, 2, 3, 4}, o2[4] = {1, 3, 2, 4}, o1[4] = {2, 3, 1, 4};
    const int *o = (l == 4 || l == 3) ? o34 : (l == 2 ? o2 : o1);
    memcpy(order, o, sizeof(o34));
}

/* old_sort_6 chooses its loop order from n (old_sort.f:111,136,161,186,211). */
static void old6_order(int n, int *order)
{
    static const int o56[6] = {1, 2, 3, 4, 5, 6}, o4[6] = {1, 2, 3, 5, 4, 6},
                     o3[6] = {1, 2, 4, 5, 3, 6}, o2[6] = {1, 3, 4, 5, 2, 6},
                     o1[6] = {2, 3, 4, 5, 1, 6};
    const int *o = (n == 6 || n == 5) ? o56 : n == 4 ? o4 : n == 3 ? o3 : n == 2 ? o2 : o1;
    memcpy(order, o, sizeof(o56));
}

/* trunk/old_sort.f:33-95 */
void oracle_old_sort_4(const double *x, double *y, int a, int b, int c, int d,
                       int i, int j, int k, int l, double s)
{
    int dims[4] = {a, b, c, d}, perm[4] = {i, j, k, l}, order[4];
    old4_order(l, order);
    walk(x, y, 4, dims, perm, order, s, 0.0, 0);
}

/* trunk/old_sort.f:512-574 */
void oracle_old_sortacc_4(const double *x, double *y, int a, int b, int c, int d,
                          int i, int j, int k, int l, double s)
{
    int dims[4] = {a, b, c, d}, perm[4] = {i, j, k, l}, order[4];
    old4_order(l, order);
    walk(x, y, 4, dims, perm, order, s, 0.0, 1);
}

/* trunk/old_sort.f:97-234 */
void oracle_old_sort_6(const double *x, double *y, int a, int b, int c, int d,
                       int e, int f, int i, int j, int k, int l, int m, int n,
                       double s)
{
    int dims[6] = {a, b, c, d, e, f}, perm[6] = {i, j, k, l, m, n}, order[6];
    old6_order(n, order);
    walk(x, y, 6, dims, perm, order, s, 0.0, 0);
}

/* trunk/old_sort.f:576-713 */
void oracle_old_sortacc_6(const double *x, double *y, int a, int b, int c, int d,
                          int e, int f, int i, int j, int k, int l, int m, int n,
                          double s)
{
    int dims[6] = {a, b, c, d, e, f}, perm[6] = {i, j, k, l, m, n}, order[6];
    old6_order(n, order);
    walk(x, y, 6, dims, perm, order, s, 0.0, 1);
}

/* Any rank 1..8, natural loop order (old_sort_2: old_sort.f:10-31,
 * old_sort_8: old_sort.f:236-478 follow the same index rule).
 * mode 0: sort, 1: sort-accumulate, 2: y = g*y + f*x (generator2.py:396). */
int oracle_sort_n(const double *x, double *y, int R, const int *dims,
                  const int *perm, double f, double g, int mode)
{
    if (R == 0) {
        /* old_sort_0 / old_sortacc_0 (old_sort.f:1-8): a scalar */
        double p = x[0] * f;
        y[0] = mode == 0 ? p : (mode == 1 ? y[0] + p : y[0] * g + p);
        return 0;
    }
    if (R < 1 || R > ORACLE_MAX_RANK) return -1;
    int seen[ORACLE_MAX_RANK] = {0}, order[ORACLE_MAX_RANK];
    for (int q = 0; q < R; ++q) {
        if (perm[q] < 1 || perm[q] > R || seen[perm[q] - 1]) return -1;
        seen[perm[q] - 1] = 1;
        if (dims[q] < 0) return -1;
        order[q] = q + 1;
    }
    walk(x, y, R, dims, perm, order, f, g, mode);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Generated library kernel, trans_<P>_loop_<L>_omp_any_f                   */
/* (generator.py:107-159; accumulate form generator.py:143-144;             */
/* 6D generator6.py:137-220).  Loop L (1-based axis numbers, outer->inner), */
/* OpenMP `parallel do collapse(2) schedule(static)` over the two outermost */
/* loops for rank 4 (generator.py:129-137) and collapse(3) for rank 6       */
/* (generator6.py:172-183).  Used as the timed CPU baseline.                */
/* mode 0: sorted = f*unsorted (generator.py:150)                           */
/* mode 1: sorted = sorted + f*unsorted (generator.py:143-144,150)          */
/* mode 2: sorted = g*sorted + f*unsorted (generator6.py:205-208)           */
/* ------------------------------------------------------------------------ */
int oracle_trans(const double *u, double *s, int R, const int *dims,
                 const int *perm, const int *loop, double f, double g,
                 int mode, int nthreads)
{
    if (R < 2 || R > ORACLE_MAX_RANK) return -1;
    int64_t istr[ORACLE_MAX_RANK], ostr[ORACLE_MAX_RANK];
    in_strides(R, dims, istr);
    out_strides(R, dims, perm, ostr);
    int collapse = (R >= 6) ? 3 : 2;
    if (collapse > R - 1) collapse = R - 1;
    int ax[ORACLE_MAX_RANK];
    for (int t = 0; t < R; ++t) ax[t] = loop[t] - 1;
    int64_t nouter = 1;
    for (int t = 0; t < collapse; ++t) nouter *= dims[ax[t]];
    int64_t ninner = 1;
    for (int t = collapse; t < R; ++t) ninner *= dims[ax[t]];
    if (nouter <= 0 || ninner <= 0) return 0;
    const int inner = ax[R - 1];
    const int64_t is = istr[inner], os = ostr[inner];
    const int nin = dims[inner];
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t o = 0; o < nouter; ++o) {
        /* coordinates of the collapsed outer loops */
        int64_t ia = 0, ib = 0, rem = o;
        for (int t = collapse - 1; t >= 0; --t) {
            int a = ax[t];
            int64_t c = rem % dims[a];
            rem /= dims[a];
            ia += c * istr[a];
            ib += c * ostr[a];
        }
        /* remaining middle loops, then the innermost */
        int idx[ORACLE_MAX_RANK] = {0};
        for (;;) {
            const double *xp = u + ia;
            double *yp = s + ib;
            if (mode == 0) {
                for (int t = 0; t < nin; ++t) yp[t * os] = f * xp[t * is];
            } else if (mode == 1) {
                for (int t = 0; t < nin; ++t) {
                    double p = f * xp[t * is];
                    yp[t * os] = yp[t * os] + p;
                }
            } else {
                for (int t = 0; t < nin; ++t) {
                    double p = f * xp[t * is];
                    double q = g * yp[t * os];
                    yp[t * os] = q + p;
                }
            }
            int lev = R - 2;
            for (; lev >= collapse; --lev) {
                int a = ax[lev];
                idx[a]++;
                ia += istr[a];
                ib += ostr[a];
                if (idx[a] < dims[a]) break;
                ia -= istr[a] * dims[a];
                ib -= ostr[a] * dims[a];
                idx[a] = 0;
            }
            if (lev < collapse) break;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Inputs.                                                                  */
/* ------------------------------------------------------------------------ */

/* init_4d_array / init_6d_array (tester_futil.F:139-185):
 * x(i,j,k,l) = i + j*a + k*a*b + l*a*b*c with the FIRST Fortran index
 * fastest.  At 0-based linear position p this is p + (1 + a + ab + abc),
 * a plain ramp; the dims argument lists the Fortran extents in the order
 * the tester passes them (dim1 fastest for this initialiser). */
void oracle_init_kat(double *x, int R, const int *dims)
{
    int64_t off = 0, prod = 1, n = 1;
    for (int t = 0; t < R; ++t) n *= dims[t];
    for (int t = 0; t < R; ++t) {
        off += prod;
        prod *= dims[t];
    }
    for (int64_t p = 0; p < n; ++p) x[p] = (double)(p + off);
}

/* splitmix64 of (seed + (i+1)*golden): a stateless, index-addressable
 * generator, so Python (NumPy) and C produce byte-identical inputs. */
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* uniform(-1,1): top 53 bits -> [0,1), then 2u-1 (both steps exact). */
void oracle_fill_uniform(double *x, int64_t n, uint64_t seed)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < n; ++i) {
        uint64_t z = mix64(seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL);
        double u = (double)(z >> 11) * 0x1.0p-53;
        x[i] = 2.0 * u - 1.0;
    }
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
