/*
 * Precision-generic body of the CPU oracle (test infrastructure only).
 * Included twice by fslbm_oracle.c with
 *   REAL = double, SFX = f64, EXACT = 1   (exact rational constants)
 *   REAL = float,  SFX = f32, EXACT = 0   (the reference's float literals)
 * Arithmetic follows the source order of FSLBM.cu and is compiled with
 * -ffp-contract=off, so every + and * rounds once, as written.
 */
#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC_FN(name) ORC_CAT(name, SFX)

/* Scalars of one launch, FSLBM.cu:231,269,599,623-624. */
typedef struct {
    REAL W[9];           /* weights of the cx=0 plane, FSLBM.cu:60-61 */
    REAL quarter;        /* cx=+-1 factor (1/6)/(2/3), FSLBM.cu:159,171 */
    REAL dtddy, m, c, cc;
    REAL c0, c1, c2;
    REAL pin;
    int d1;
} ORC_FN(orc_k);

static void ORC_FN(orc_setup)(const orc_params *prm, ORC_FN(orc_k) * k)
{
#if EXACT
    const double w0 = 8.0 / 27.0, w1 = 2.0 / 27.0, w2 = 1.0 / 54.0;
    double b = prm->blend;
    k->dtddy = prm->nu - 1.0 / 6.0;
    k->m = 1.0 / (double)(prm->d1 * prm->d1);
    k->c = prm->tblend;
    k->cc = 1.0 - prm->tblend;
    k->c0 = -6.0 - 1.5 * b;
    k->c1 = 1.0 + b / 3.0;
    k->c2 = -b / 12.0;
    k->pin = prm->pin;
#else
    /* FSLBM.cu:60-61 literals; tblend/pin/nu/blend are floats read by %e */
    const float w0 = 0.296296296f, w1 = 0.07407407407f, w2 = 0.01851851851f;
    float tb = (float)prm->tblend, nu = (float)prm->nu, b = (float)prm->blend;
    k->dtddy = nu - 0.16666666f;                  /* FSLBM.cu:599 */
    k->m = (float)(1.0 / (double)(float)(prm->d1 * prm->d1)); /* :624 */
    k->c = tb;                                    /* :269 */
    k->cc = (float)(1.0 - (double)tb);            /* :269, double then float */
    k->c0 = -6.0f - 1.5f * b;                     /* :231 */
    k->c1 = 1.0f + 0.3333333333f * b;
    k->c2 = -0.0833333333f * b;
    k->pin = (float)prm->pin;
#endif
    k->W[0] = (REAL)w0;
    for (int i = 1; i <= 4; i++) k->W[i] = (REAL)w1;
    for (int i = 5; i <= 8; i++) k->W[i] = (REAL)w2;
    k->quarter = (REAL)0.25;
    k->d1 = prm->d1;
}

/* Fetch of one neighbour for the predictor: value and wall flag. */
typedef struct {
    REAL p, u, v, w;
    int wall;
} ORC_FN(orc_nb);

/* Predictor + corrector + blend for one fluid cell, FSLBM.cu:130-273.
 * nb[(cz+1)*9 + (cy+1)*3 + (cx+1)] holds the 27 gathered neighbours
 * (pressure already carrying any periodic jump); fd[axis][j] holds the
 * velocity triples at x + s*e_axis for s = +d1, -d1, +2d1, -2d1. */
static void ORC_FN(orc_cell)(const ORC_FN(orc_k) * k,
                             const ORC_FN(orc_nb) * nb,
                             const REAL fd[3][4][3], REAL out[4])
{
    static const int cy[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
    static const int cz[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
    const ORC_FN(orc_nb) *self = &nb[13];
    REAL rhd1 = self->p, u = self->u, v = self->v, w = self->w;
    REAL rt1 = 0, uxt1v = 0, uyt1v = 0, uzt1v = 0;
    for (int i = 8; i >= 0; i--) {
        REAL fy = (REAL)cy[i], fz = (REAL)cz[i];
        int base = (cz[i] + 1) * 9 + (cy[i] + 1) * 3;
        /* cx = 0, FSLBM.cu:137-150 */
        {
            const ORC_FN(orc_nb) *n = &nb[base + 1];
            REAL px = n->wall ? rhd1 : n->p;
            REAL zet = (fy * n->v + fz * n->w) * (REAL)3;
            REAL uu = (n->u * n->u + n->v * n->v + n->w * n->w) * (REAL)3;
            REAL f = k->W[i] * (px + zet + (REAL)0.5 * (zet * zet - uu));
            rt1 += f;
            uyt1v += f * fy;
            uzt1v += f * fz;
        }
        /* cx = +1, FSLBM.cu:152-161 */
        {
            const ORC_FN(orc_nb) *n = &nb[base + 2];
            REAL px = n->wall ? rhd1 : n->p;
            REAL zet = (n->u + fy * n->v + fz * n->w) * (REAL)3;
            REAL uu = (n->u * n->u + n->v * n->v + n->w * n->w) * (REAL)3;
            REAL f = k->W[i] * (px + zet + (REAL)0.5 * (zet * zet - uu)) *
                     k->quarter;
            rt1 += f;
            uyt1v += f * fy;
            uxt1v += f;
            uzt1v += f * fz;
        }
        /* cx = -1, FSLBM.cu:164-173 */
        {
            const ORC_FN(orc_nb) *n = &nb[base + 0];
            REAL px = n->wall ? rhd1 : n->p;
            REAL zet = (-n->u + fy * n->v + fz * n->w) * (REAL)3;
            REAL uu = (n->u * n->u + n->v * n->v + n->w * n->w) * (REAL)3;
            REAL f = k->W[i] * (px + zet + (REAL)0.5 * (zet * zet - uu)) *
                     k->quarter;
            rt1 += f;
            uyt1v += f * fy;
            uxt1v -= f;
            uzt1v += f * fz;
        }
    }
    /* CCS corrector, FSLBM.cu:236-257; fd order matches uvw[0..5]:
     * +x, -x, +y, -y, +z, -z at distance d1 then 2*d1. */
    REAL L[3];
    const REAL own[3] = {u, v, w};
    for (int a = 0; a < 3; a++) {
        REAL t = own[a] * k->c0;
        REAL s1 = fd[0][0][a] + fd[0][1][a] + fd[1][0][a] + fd[1][1][a] +
                  fd[2][0][a] + fd[2][1][a];
        t += k->c1 * s1;
        REAL s2 = fd[0][2][a] + fd[0][3][a] + fd[1][2][a] + fd[1][3][a] +
                  fd[2][2][a] + fd[2][3][a];
        t += k->c2 * s2;
        L[a] = t;
    }
    /* blend, FSLBM.cu:269-273 */
    out[0] = rt1;
    out[1] = (uxt1v + k->dtddy * L[0] * k->m) * k->c + u * k->cc;
    out[2] = (uyt1v + k->dtddy * L[1] * k->m) * k->c + v * k->cc;
    out[3] = (uzt1v + k->dtddy * L[2] * k->m) * k->c + w * k->cc;
}

static inline int ORC_FN(orc_clamp)(int v, int n)
{
    return v < 0 ? 0 : (v >= n ? n - 1 : v);
}

/* ------------------------------------------------------------------ */
/* Padded (reference) layout                                           */
/* ------------------------------------------------------------------ */
void ORC_FN(orc_step_padded)(int lx, int ly, int lz, int d2, const uint8_t *IO,
                             const orc_params *prm, double p2d,
                             const REAL *in, REAL *out)
{
    ORC_FN(orc_k) k;
    ORC_FN(orc_setup)(prm, &k);
    const long N = (long)lx * ly * lz, lxy = (long)lx * ly;
    const REAL *P = in, *U = in + N, *V = in + 2 * N, *Wv = in + 3 * N;
    const REAL p2 = (REAL)p2d;
    REAL *pre = (REAL *)malloc(sizeof(REAL) * 4 * N);
    /* Pass 1: FSLBMcalc on every IO==0 cell (texture reads clamp to the
     * volume, FSLBM.cu:85-123,243-253); other cells keep time level n. */
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < lz; z++)
        for (int y = 0; y < ly; y++)
            for (int x = 0; x < lx; x++) {
                long idx = x + lx * (long)y + lxy * z;
                if (IO[idx] != 0) {
                    for (int f = 0; f < 4; f++) pre[f * N + idx] = in[f * N + idx];
                    continue;
                }
                ORC_FN(orc_nb) nb[27];
                for (int cz = -1; cz <= 1; cz++)
                    for (int cy = -1; cy <= 1; cy++)
                        for (int cx = -1; cx <= 1; cx++) {
                            int xs = x + cx, ys = y + cy, zs = z + cz;
                            ORC_FN(orc_nb) *n = &nb[(cz + 1) * 9 + (cy + 1) * 3 + (cx + 1)];
                            int inside = xs >= 0 && xs < lx && ys >= 0 && ys < ly &&
                                         zs >= 0 && zs < lz;
                            long j = ORC_FN(orc_clamp)(xs, lx) +
                                     lx * (long)ORC_FN(orc_clamp)(ys, ly) +
                                     lxy * ORC_FN(orc_clamp)(zs, lz);
                            n->p = P[j]; n->u = U[j]; n->v = V[j]; n->w = Wv[j];
                            /* IO[idx+id2] < 125, FSLBM.cu:138; outside the
                             * volume counts as wall (unreachable with a frame) */
                            n->wall = inside ? !(IO[j] < 125) : 1;
                        }
                REAL fd[3][4][3];
                const int s[4] = {k.d1, -k.d1, 2 * k.d1, -2 * k.d1};
                for (int a = 0; a < 3; a++)
                    for (int q = 0; q < 4; q++) {
                        int xs = x, ys = y, zs = z;
                        if (a == 0) xs += s[q];
                        if (a == 1) ys += s[q];
                        if (a == 2) zs += s[q];
                        long j = ORC_FN(orc_clamp)(xs, lx) +
                                 lx * (long)ORC_FN(orc_clamp)(ys, ly) +
                                 lxy * ORC_FN(orc_clamp)(zs, lz);
                        fd[a][q][0] = U[j]; fd[a][q][1] = V[j]; fd[a][q][2] = Wv[j];
                    }
                REAL o[4];
                ORC_FN(orc_cell)(&k, nb, (const REAL(*)[4][3])fd, o);
                for (int f = 0; f < 4; f++) pre[f * N + idx] = o[f];
            }
    /* Pass 2: bcond, FSLBM.cu:280-303, reading pre-bcond values only. */
    const int L = lz - 2 * d2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < lz; z++)
        for (int y = 0; y < ly; y++)
            for (int x = 0; x < lx; x++) {
                long idx = x + lx * (long)y + lxy * z;
                if (IO[idx] == 125) {
                    for (int f = 0; f < 4; f++) out[f * N + idx] = 0;
                    continue;
                }
                if (z < d2) {
                    long j = idx + lxy * L;  /* plane lz-2*d2+z */
                    REAL pv = pre[j];
                    pv -= p2;
                    pv += k.pin;
                    out[idx] = pv;
                    for (int f = 1; f < 4; f++) out[f * N + idx] = pre[f * N + j];
                } else if (z >= lz - d2) {
                    long j = idx - lxy * L;  /* plane 2*d2-lz+z */
                    REAL pv = pre[j];
                    pv -= k.pin;
                    pv -= p2;
                    out[idx] = pv;
                    for (int f = 1; f < 4; f++) out[f * N + idx] = pre[f * N + j];
                } else {
                    out[idx] = pre[idx] - p2;
                    for (int f = 1; f < 4; f++) out[f * N + idx] = pre[f * N + idx];
                }
            }
    free(pre);
}

/* ------------------------------------------------------------------ */
/* Folded layout                                                       */
/* ------------------------------------------------------------------ */
static void ORC_FN(orc_folded_cell)(int nx, int ny, int nz, const uint8_t *mask,
                                    const ORC_FN(orc_k) * k, REAL gauge,
                                    const REAL *in, REAL *out, int x, int y,
                                    int z)
{
    const long N = (long)nx * ny * nz, nxy = (long)nx * ny;
    const REAL *P = in, *U = in + N, *V = in + 2 * N, *Wv = in + 3 * N;
    long idx = x + nx * (long)y + nxy * z;
    if (mask && mask[idx] != 0) {
        if (mask[idx] >= 125)
            for (int f = 0; f < 4; f++) out[f * N + idx] = 0;
        else
            for (int f = 0; f < 4; f++) out[f * N + idx] = in[f * N + idx];
        return;
    }
    ORC_FN(orc_nb) nb[27];
    for (int cz = -1; cz <= 1; cz++)
        for (int cy = -1; cy <= 1; cy++)
            for (int cx = -1; cx <= 1; cx++) {
                int xs = x + cx, ys = y + cy, zs = z + cz;
                REAL jump = 0;
                int wrapped = 0;
                if (zs < 0) { zs += nz; jump = k->pin; wrapped = 1; }
                else if (zs >= nz) { zs -= nz; jump = -k->pin; wrapped = 1; }
                int inside = 1;
                if (mask) {
                    inside = xs >= 0 && xs < nx && ys >= 0 && ys < ny;
                    xs = ORC_FN(orc_clamp)(xs, nx);
                    ys = ORC_FN(orc_clamp)(ys, ny);
                } else {
                    xs = (xs + nx) % nx;
                    ys = (ys + ny) % ny;
                }
                long j = xs + nx * (long)ys + nxy * zs;
                ORC_FN(orc_nb) *n = &nb[(cz + 1) * 9 + (cy + 1) * 3 + (cx + 1)];
                n->p = wrapped ? P[j] + jump : P[j];
                n->u = U[j]; n->v = V[j]; n->w = Wv[j];
                n->wall = mask ? (!inside || mask[j] >= 125) : 0;
            }
    REAL fd[3][4][3];
    const int s[4] = {k->d1, -k->d1, 2 * k->d1, -2 * k->d1};
    for (int a = 0; a < 3; a++)
        for (int q = 0; q < 4; q++) {
            int xs = x, ys = y, zs = z;
            if (a == 0) xs += s[q];
            if (a == 1) ys += s[q];
            if (a == 2) zs += s[q];
            zs = ((zs % nz) + nz) % nz;
            if (mask) {
                xs = ORC_FN(orc_clamp)(xs, nx);
                ys = ORC_FN(orc_clamp)(ys, ny);
            } else {
                xs = ((xs % nx) + nx) % nx;
                ys = ((ys % ny) + ny) % ny;
            }
            long j = xs + nx * (long)ys + nxy * zs;
            fd[a][q][0] = U[j]; fd[a][q][1] = V[j]; fd[a][q][2] = Wv[j];
        }
    REAL o[4];
    ORC_FN(orc_cell)(k, nb, (const REAL(*)[4][3])fd, o);
    out[idx] = o[0] - gauge;
    for (int f = 1; f < 4; f++) out[f * N + idx] = o[f];
}

void ORC_FN(orc_step_folded)(int nx, int ny, int nz, const uint8_t *mask,
                             const orc_params *prm, double gauge,
                             const REAL *in, REAL *out)
{
    ORC_FN(orc_k) k;
    ORC_FN(orc_setup)(prm, &k);
    const REAL g = (REAL)gauge;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                ORC_FN(orc_folded_cell)(nx, ny, nz, mask, &k, g, in, out, x, y, z);
}

double ORC_FN(orc_time_folded)(int nx, int ny, int nz, const uint8_t *mask,
                               const orc_params *prm, int nsteps, REAL *a,
                               REAL *b, int *threads)
{
    *threads = omp_get_max_threads();
    double t0 = omp_get_wtime();
    for (int s = 0; s < nsteps; s++) {
        if (s & 1) ORC_FN(orc_step_folded)(nx, ny, nz, mask, prm, 0.0, b, a);
        else ORC_FN(orc_step_folded)(nx, ny, nz, mask, prm, 0.0, a, b);
    }
    return omp_get_wtime() - t0;
}

#undef ORC_FN
#undef ORC_CAT
#undef ORC_CAT2
