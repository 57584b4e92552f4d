/*
 * CPU oracle for the FSLBM time step -- TEST INFRASTRUCTURE ONLY.
 * See fslbm_oracle.h for what each function restates and where.
 */
#include "fslbm_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define REAL double
#define SFX f64
#define EXACT 1
#include "fslbm_oracle_impl.h"
#undef REAL
#undef SFX
#undef EXACT

#define REAL float
#define SFX f32
#define EXACT 0
#include "fslbm_oracle_impl.h"
#undef REAL
#undef SFX
#undef EXACT

int orc_max_threads(void) { return omp_get_max_threads(); }

/* FSLBM.cu:632-647: q = sum of uz, p = sum of p over the whole XY plane
 * (walls hold 0), outlet plane z=lz-d2 and inlet plane z=d2, divided by
 * the cross-section fluid count l1. */
void orc_monitor_padded_f64(int lx, int ly, int lz, int d2, double l1,
                            const double *state, double out4[4])
{
    const long N = (long)lx * ly * lz, lxy = (long)lx * ly;
    const double *P = state, *W = state + 3 * N;
    double q1 = 0, q2 = 0, p1 = 0, p2 = 0;
    for (long i = 0; i < lxy; i++) {
        p2 += P[i + lxy * (lz - d2)];
        q2 += W[i + lxy * (lz - d2)];
    }
    for (long i = 0; i < lxy; i++) {
        p1 += P[i + lxy * d2];
        q1 += W[i + lxy * d2];
    }
    out4[0] = q1 / l1;
    out4[1] = q2 / l1;
    out4[2] = p1 / l1;
    out4[3] = p2 / l1;
}

/* FSLBM.cu:683-748 in the reference's order of operations. */
void orc_stats_accumulate(int n, int delta, double *acc, const double *sample,
                          int *cnt, int *cnt2)
{
    double *pav = acc, *uxav = acc + n, *uyav = acc + 2 * (long)n,
           *uzav = acc + 3 * (long)n, *p2av = acc + 4 * (long)n,
           *ux2av = acc + 5 * (long)n, *uy2av = acc + 6 * (long)n,
           *uz2av = acc + 7 * (long)n, *uxyav = acc + 8 * (long)n,
           *uxzav = acc + 9 * (long)n, *uyzav = acc + 10 * (long)n;
    const double *sp = sample, *sx = sample + n, *sy = sample + 2 * (long)n,
                 *sz = sample + 3 * (long)n;
    if (delta == 0) {
        *cnt = 0;
        *cnt2 = 0;
        memset(acc, 0, sizeof(double) * 11 * (size_t)n);
    }
    if (delta == 1) {
        *cnt2 = 0;
        memset(acc + 4 * (long)n, 0, sizeof(double) * 7 * (size_t)n);
    }
    *cnt += 1;
    *cnt2 += 1;
    const double c = (double)*cnt;
    for (long k = 0; k < n; k++) {
        pav[k] = pav[k] + sp[k];
        double dd1 = pav[k] / c - sp[k];
        uxav[k] = uxav[k] + sx[k];
        double dd2 = uxav[k] / c - sx[k];
        uyav[k] = uyav[k] + sy[k];
        double dd3 = uyav[k] / c - sy[k];
        uzav[k] = uzav[k] + sz[k];
        double dd4 = uzav[k] / c - sz[k];
        p2av[k] = p2av[k] + fabs(dd1);
        ux2av[k] = ux2av[k] + fabs(dd2);
        uy2av[k] = uy2av[k] + fabs(dd3);
        uz2av[k] = uz2av[k] + fabs(dd4);
        uxyav[k] = uxyav[k] + dd2 * dd3;
        uxzav[k] = uxzav[k] + dd2 * dd4;
        uyzav[k] = uyzav[k] + dd4 * dd3;
    }
}
