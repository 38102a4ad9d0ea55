/*
 * CPU oracle for the FFT docking rotation loop -- TEST INFRASTRUCTURE ONLY.
 * See mdk_oracle.h.  Built with -O3 -ffp-contract=off (no FMA contraction), like the
 * reference's x86-64 build without -march (Makefile:15), so float expressions round
 * after every operation exactly as there.
 */
#include "mdk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int or_version(void) { return 1; }

/* ------------------------------------------------------------------ A2 */
/* src/docking.cpp:376-384.  The reference calls cos()/sin() on float arguments in
 * C++, which resolves to the float overloads (cosf/sinf); products are float. */
void or_rotation_coef(const float theta[3], float r[9])
{
    const float t0 = theta[0], t1 = theta[1], t2 = theta[2];
    r[0] = cosf(t0) * cosf(t2) - sinf(t0) * cosf(t1) * sinf(t2);  /* r11 */
    r[3] = sinf(t0) * cosf(t2) + cosf(t0) * cosf(t1) * sinf(t2);  /* r21 */
    r[6] = sinf(t1) * sinf(t2);                                    /* r31 */
    r[1] = -cosf(t0) * sinf(t2) - sinf(t0) * cosf(t1) * cosf(t2); /* r12 */
    r[4] = -sinf(t0) * sinf(t2) + cosf(t0) * cosf(t1) * cosf(t2); /* r22 */
    r[7] = sinf(t1) * cosf(t2);                                    /* r32 */
    r[2] = sinf(t0) * sinf(t1);                                    /* r13 */
    r[5] = -cosf(t0) * sinf(t1);                                   /* r23 */
    r[8] = cosf(t1);                                               /* r33 */
}

/* src/docking.cpp:386-393 */
void or_rotate_ligand(int na, const float *coord, const float center[3], const float r[9],
                      float *out)
{
    for (int i = 0; i < na; i++) {
        const float x = coord[3 * i] - center[0];
        const float y = coord[3 * i + 1] - center[1];
        const float z = coord[3 * i + 2] - center[2];
        out[3 * i] = r[0] * x + r[1] * y + r[2] * z;
        out[3 * i + 1] = r[3] * x + r[4] * y + r[5] * z;
        out[3 * i + 2] = r[6] * x + r[7] * y + r[8] * z;
    }
}

/* ------------------------------------------------------------------ A3 */
/* src/docking.cpp:133-141: float search_length; the expression is evaluated in double
 * and rounded once to float. */
void or_grid_coord(int ng, float gw, float *out)
{
    const float search_length = gw * (float)ng;
    for (int i = 0; i < ng; i++)
        out[i] = (float)((double)search_length * (-0.5 + (0.5 + i) / (double)ng));
}

/* src/protein.cpp:67-111 (float edges; centre = float sum / 2.0 in double, stored float) */
void or_center(int na, const float *coord, float center[3])
{
    float lo[3] = {1.0e9f, 1.0e9f, 1.0e9f}, hi[3] = {-1.0e9f, -1.0e9f, -1.0e9f};
    for (int i = 0; i < na; i++)
        for (int j = 0; j < 3; j++) {
            const float c = coord[3 * i + j];
            if (c < lo[j]) lo[j] = c;
            if (c > hi[j]) hi[j] = c;
        }
    for (int j = 0; j < 3; j++) center[j] = (float)((double)(lo[j] + hi[j]) / 2.0);
}

/* src/protein.cpp:114-131 */
void or_shift_center(int na, const float *coord, float *out)
{
    float c[3];
    or_center(na, coord, c);
    for (int i = 0; i < na; i++)
        for (int j = 0; j < 3; j++) out[3 * i + j] = coord[3 * i + j] - c[j];
}

/* src/protein.cpp:134-165: xd[l*ng+i] = (x_l - g_i)^2 in float. */
static void voxel_dist(int ng, int na, const float *xyz, const float *gc, float *xd, float *yd,
                       float *zd)
{
    for (int l = 0, k = 0; l < na; l++)
        for (int i = 0; i < ng; i++, k++) {
            float dx = xyz[3 * l] - gc[i], dy = xyz[3 * l + 1] - gc[i], dz = xyz[3 * l + 2] - gc[i];
            xd[k] = dx * dx;
            yd[k] = dy * dy;
            zd[k] = dz * dz;
        }
}

static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ A4 + A5 */
/* src/ligand.cpp:20-197 (shape), :200-239 (charge), :242-261 (radii^2). */
void or_ligand_voxel(int ng, float gw, int na, const float *xyz, const float *radius,
                     const float *charge, const float *gc, float delta, float *grid_r,
                     float *grid_i)
{
    const float surface = 1.0f, swollen = -8888.0f;
    const int ng2 = ng * ng, ng3 = ng2 * ng;
    float *xd = (float *)malloc(sizeof(float) * (size_t)na * ng * 3);
    float *yd = xd + (size_t)na * ng, *zd = yd + (size_t)na * ng;
    float *rc2 = (float *)malloc(sizeof(float) * (size_t)na * 2), *rs2 = rc2 + na;
    for (int i = 0; i < na; i++) {  /* src/ligand.cpp:254-257 */
        rc2[i] = radius[i] * radius[i] * 1.5f;
        rs2[i] = radius[i] * radius[i] * 1.0f;
    }
    memset(grid_r, 0, sizeof(float) * ng3);
    memset(grid_i, 0, sizeof(float) * ng3);
    voxel_dist(ng, na, xyz, gc, xd, yd, zd);

    for (int l = 0; l < na; l++) { /* core fill, src/ligand.cpp:68-101 */
        const int sr = 2, lc = ng * l;
        const int i2 = (int)(xyz[3 * l] / gw + ng / 2);
        const int j2 = (int)(xyz[3 * l + 1] / gw + ng / 2);
        const int k2 = (int)(xyz[3 * l + 2] / gw + ng / 2);
        const int ia = imax(i2 - sr, 0), ja = imax(j2 - sr, 0), ka = imax(k2 - sr, 0);
        const int ib = imin(i2 + sr + 1, ng), jb = imin(j2 + sr + 1, ng), kb = imin(k2 + sr + 1, ng);
        for (int i = ia; i < ib; i++) {
            if (xd[lc + i] > rc2[l]) continue;
            for (int j = ja; j < jb; j++) {
                const float d2 = xd[lc + i] + yd[lc + j];
                if (d2 > rc2[l]) continue;
                for (int k = ka; k < kb; k++) {
                    const float d3 = d2 + zd[lc + k];
                    if (d3 < rc2[l]) grid_r[ng2 * i + ng * j + k] = delta;
                }
            }
        }
    }
    /* scrape swollen surface, src/ligand.cpp:110-134 */
    for (int i = 1; i < ng - 1; i++)
        for (int j = 1; j < ng - 1; j++)
            for (int k = 1, ijk = ng2 * i + ng * j + 1; k < ng - 1; k++, ijk++)
                if (grid_r[ijk] == delta &&
                    (grid_r[ijk - 1] == 0 || grid_r[ijk + 1] == 0 || grid_r[ijk - ng] == 0 ||
                     grid_r[ijk + ng] == 0 || grid_r[ijk - ng2] == 0 || grid_r[ijk + ng2] == 0))
                    grid_r[ijk] = swollen;
    for (int ijk = ng2; ijk < ng3 - ng2; ijk++)
        if (grid_r[ijk] == swollen) grid_r[ijk] = 0.0f;

    for (int l = 0; l < na; l++) { /* surface, src/ligand.cpp:140-173 */
        const int sr = 2, lc = ng * l;
        const int i2 = (int)(xyz[3 * l] / gw + ng / 2);
        const int j2 = (int)(xyz[3 * l + 1] / gw + ng / 2);
        const int k2 = (int)(xyz[3 * l + 2] / gw + ng / 2);
        const int ia = imax(i2 - sr, 0), ja = imax(j2 - sr, 0), ka = imax(k2 - sr, 0);
        const int ib = imin(i2 + sr + 1, ng), jb = imin(j2 + sr + 1, ng), kb = imin(k2 + sr + 1, ng);
        for (int i = ia; i < ib; i++) {
            if (xd[lc + i] > rs2[l]) continue;
            for (int j = ja; j < jb; j++) {
                const float d2 = xd[lc + i] + yd[lc + j];
                if (d2 > rs2[l]) continue;
                for (int k = ka; k < kb; k++) {
                    const int ijk = ng2 * i + ng * j + k;
                    if (grid_r[ijk] == delta) continue;
                    const float d3 = d2 + zd[lc + k];
                    if (d3 < rs2[l]) grid_r[ijk] = surface;
                }
            }
        }
    }
    if (surface != delta) { /* src/ligand.cpp:176-195 */
        for (int ijk = ng2; ijk < ng3 - ng2; ijk++) {
            const int i = ijk / ng2, j = (ijk / ng) % ng, k = ijk % ng;
            if (i == 0 || i == ng - 1 || j == 0 || j == ng - 1 || k == 0 || k == ng - 1) continue;
            if (grid_r[ijk] == delta &&
                (grid_r[ijk - 1] == 0 || grid_r[ijk + 1] == 0 || grid_r[ijk - ng] == 0 ||
                 grid_r[ijk + ng] == 0 || grid_r[ijk - ng2] == 0 || grid_r[ijk + ng2] == 0))
                grid_r[ijk] = surface;
        }
    }
    /* charges, src/ligand.cpp:212-235 (atom order, nearest voxel, clamped) */
    const float pad = (float)ng * gw / 2;
    for (int l = 0; l < na; l++) {
        const int i = imax(0, imin(ng - 1, (int)((xyz[3 * l] + pad) / gw)));
        const int j = imax(0, imin(ng - 1, (int)((xyz[3 * l + 1] + pad) / gw)));
        const int k = imax(0, imin(ng - 1, (int)((xyz[3 * l + 2] + pad) / gw)));
        grid_i[i * ng2 + j * ng + k] += charge[l];
    }
    free(xd);
    free(rc2);
}

/* ------------------------------------------------------------------ A14 */
/* src/receptor.cpp:20-254 (rpscace), :258-346 (electro), :349-372 (precalc). */
void or_receptor_voxel(int ng, float gw, int na, const float *xyz, const float *radius,
                       const float *charge, const float *ace, const float *gc, float aceratio,
                       float rec_core, float eratio, float *grid_r, float *grid_i)
{
    const float epsilon = rec_core, open_space = -7777.0f;
    const int ng2 = ng * ng, ng3 = ng2 * ng;
    float *vox = (float *)calloc((size_t)ng3, sizeof(float));
    float *xd = (float *)malloc(sizeof(float) * (size_t)na * ng * 3);
    float *yd = xd + (size_t)na * ng, *zd = yd + (size_t)na * ng;
    float *rj2 = (float *)malloc(sizeof(float) * (size_t)na * 2), *rc2 = rj2 + na;
    for (int i = 0; i < na; i++) { /* src/receptor.cpp:362-366 */
        rj2[i] = (radius[i] + 3.6f) * (radius[i] + 3.6f);
        rc2[i] = radius[i] * radius[i] * 1.5f;
    }
    memset(grid_r, 0, sizeof(float) * ng3);
    memset(grid_i, 0, sizeof(float) * ng3);
    voxel_dist(ng, na, xyz, gc, xd, yd, zd);

    for (int l = 0; l < na; l++) { /* core, src/receptor.cpp:70-101 */
        const int sr = (int)((2.4 + (double)gw - 0.01) / (double)gw), lc = ng * l;
        const int i2 = (int)(xyz[3 * l] / gw + ng / 2);
        const int j2 = (int)(xyz[3 * l + 1] / gw + ng / 2);
        const int k2 = (int)(xyz[3 * l + 2] / gw + ng / 2);
        const int ia = imax(i2 - sr, 0), ja = imax(j2 - sr, 0), ka = imax(k2 - sr, 0);
        const int ib = imin(i2 + sr + 1, ng), jb = imin(j2 + sr + 1, ng), kb = imin(k2 + sr + 1, ng);
        for (int i = ia; i < ib; i++) {
            if (xd[lc + i] > rc2[l]) continue;
            for (int j = ja; j < jb; j++) {
                const float d2 = xd[lc + i] + yd[lc + j];
                if (d2 > rc2[l]) continue;
                for (int k = ka; k < kb; k++) {
                    const float d3 = xd[lc + i] + yd[lc + j] + zd[lc + k];
                    if (d3 < rc2[l]) grid_r[ng2 * i + ng * j + k] = epsilon;
                }
            }
        }
    }
    /* cavity fill phase 1/2: six ray sweeps, src/receptor.cpp:106-157 */
    for (int ij = 0; ij < ng2; ij++) {
        const int i = ij % ng, j = ij / ng;
        for (int k = 0; k < ng; k++) {
            float *g = &grid_r[i + ng * j + ng2 * k];
            if (*g != epsilon) *g = open_space; else break;
        }
        for (int k = ng - 1; k >= 0; k--) {
            float *g = &grid_r[i + ng * j + ng2 * k];
            if (*g != epsilon) *g = open_space; else break;
        }
        for (int k = 0; k < ng; k++) {
            float *g = &grid_r[k + ng * i + ng2 * j];
            if (*g != epsilon) *g = open_space; else break;
        }
        for (int k = ng - 1; k >= 0; k--) {
            float *g = &grid_r[k + ng * i + ng2 * j];
            if (*g != epsilon) *g = open_space; else break;
        }
        for (int k = 0; k < ng; k++) {
            float *g = &grid_r[j + ng * k + ng2 * i];
            if (*g != epsilon) *g = open_space; else break;
        }
        for (int k = ng - 1; k >= 0; k--) {
            float *g = &grid_r[j + ng * k + ng2 * i];
            if (*g != epsilon) *g = open_space; else break;
        }
    }
    /* phase 2/2: two in-place, order-dependent propagation sweeps, src/receptor.cpp:159-189 */
    for (int ink = 0; ink < 2; ink++)
        for (int ijk = ng2; ijk < ng3 - ng2; ijk++) {
            const int j = (ijk / ng) % ng;
            if (j < 2 || j > ng - 2) continue;
            const int k = ijk % ng;
            if (k < 2 || k > ng - 2) continue;
            if (grid_r[ijk] == open_space) {
                if (grid_r[ijk + 1] == 0) grid_r[ijk + 1] = open_space;
                if (grid_r[ijk - 1] == 0) grid_r[ijk - 1] = open_space;
                if (grid_r[ijk + ng] == 0) grid_r[ijk + ng] = open_space;
                if (grid_r[ijk - ng] == 0) grid_r[ijk - ng] = open_space;
                if (grid_r[ijk + ng2] == 0) grid_r[ijk + ng2] = open_space;
                if (grid_r[ijk - ng2] == 0) grid_r[ijk - ng2] = open_space;
            }
        }
    for (int i = 0; i < ng3; i++) if (grid_r[i] == 0) grid_r[i] = epsilon;       /* :192-196 */
    for (int i = 0; i < ng3; i++) if (grid_r[i] == open_space) grid_r[i] = 0;    /* :198-202 */

    for (int l = 0; l < na; l++) { /* surface rPSC + ACE, src/receptor.cpp:209-243 */
        const int sr = (int)((12 + (double)gw - 0.01) / (double)gw), lc = ng * l;
        const int i2 = (int)(xyz[3 * l] / gw + ng / 2);
        const int j2 = (int)(xyz[3 * l + 1] / gw + ng / 2);
        const int k2 = (int)(xyz[3 * l + 2] / gw + ng / 2);
        const int ia = imax(i2 - sr, 0), ja = imax(j2 - sr, 0), ka = imax(k2 - sr, 0);
        const int ib = imin(i2 + sr + 1, ng), jb = imin(j2 + sr + 1, ng), kb = imin(k2 + sr + 1, ng);
        const double add = 1.0 + (double)ace[l] * (-0.8) * (double)aceratio;
        for (int i = ia; i < ib; i++) {
            if (xd[lc + i] > rj2[l]) continue;
            for (int j = ja; j < jb; j++) {
                const float d2 = xd[lc + i] + yd[lc + j];
                if (d2 > rj2[l]) continue;
                for (int k = ka; k < kb; k++) {
                    const int ijk = ng2 * i + ng * j + k;
                    const float d3 = xd[lc + i] + yd[lc + j] + zd[lc + k];
                    if (d3 < rj2[l]) vox[ijk] = (float)((double)vox[ijk] + add);
                }
            }
        }
    }
    for (int ijk = 0; ijk < ng3; ijk++) /* :245-249 */
        if (grid_r[ijk] != epsilon) grid_r[ijk] = vox[ijk];

    /* electrostatics, src/receptor.cpp:258-346 (beta = -2800, src/docking.cpp:176) */
    {
        const float beta = -2800.0f;
        const float gwl = gc[1] - gc[0];
        const float cpg = -1 * gwl * 80.0f / (beta * eratio);
        for (int l = 0; l < na; l++) {
            const float ac = fabsf(charge[l]);
            if (ac <= 1.0e-6f) continue;
            int sr;
            if ((double)ac < 0.07) sr = 4;
            else if ((double)ac < 0.21) sr = 5;
            else sr = (int)(ac / cpg);
            const int lc = ng * l;
            const int i2 = (int)(xyz[3 * l] / gwl + ng / 2);
            const int j2 = (int)(xyz[3 * l + 1] / gwl + ng / 2);
            const int k2 = (int)(xyz[3 * l + 2] / gwl + ng / 2);
            const int ia = imax(i2 - sr, 0), ja = imax(j2 - sr, 0), ka = imax(k2 - sr, 0);
            const int ib = imin(i2 + sr + 1, ng), jb = imin(j2 + sr + 1, ng), kb = imin(k2 + sr + 1, ng);
            for (int i = ia; i < ib; i++)
                for (int j = ja; j < jb; j++) {
                    const int ijkb = ng2 * i + ng * j;
                    const float d2 = xd[lc + i] + yd[lc + j];
                    for (int k = ka; k < kb; k++) {
                        const int ijk = ijkb + k;
                        if (grid_r[ijk] < 0.0) continue;
                        float d3 = d2 + zd[lc + k];
                        d3 = sqrtf(d3);
                        float er;
                        if (d3 >= 8.0f) er = 80.0f;
                        else if (d3 > 6.0f) er = 38.0f * d3 + -224.0f;
                        else er = 4.0f;
                        const float elec = eratio * beta * charge[l] / er / d3;
                        grid_i[ijk] += elec;
                    }
                }
        }
    }
    free(vox);
    free(xd);
    free(rj2);
}

/* ------------------------------------------------------------------ A7 */
void or_embed(int ng, int nf, const float *grid_r, const float *grid_i, float *c)
{
    const size_t nf3 = (size_t)nf * nf * nf;
    const int nd = (nf - ng) / 2;
    memset(c, 0, sizeof(float) * 2 * nf3);
    for (int i = 0, m = 0; i < ng; i++)
        for (int j = 0; j < ng; j++)
            for (int k = 0; k < ng; k++, m++) {
                const size_t p = ((size_t)(i + nd) * nf + (j + nd)) * nf + (k + nd);
                c[2 * p] = grid_r[m];
                c[2 * p + 1] = grid_i[m];
            }
}

/* ------------------------------------------------------------------ FFT */
/* Mixed-radix Stockham autosort FFT in float with twiddles rounded from double.
 * Stage with radix R and span Ns: for j < n/R, k = j mod Ns,
 *   v_r = x[j + r n/R] * w^(r k), w = exp(sign 2 pi i / (Ns R)),
 *   y[(j/Ns) Ns R + k + m Ns] = sum_r v_r exp(sign 2 pi i r m / R). */
static void dft_table(int R);

typedef struct {
    int n, nfac, fac[32];
    float *tw; /* per stage: (R-1)*Ns complex twiddles, concatenated */
    float *dft; /* per distinct radix: R*R complex small-DFT matrix (forward) */
} fft_plan;

static int factorize(int n, int *fac)
{
    int nf = 0;
    static const int pref[] = {4, 2, 3, 5, 7, 11, 13};
    for (int a = 0; a < 7 && n > 1; a++)
        while (n % pref[a] == 0) { fac[nf++] = pref[a]; n /= pref[a]; }
    for (int p = 17; n > 1; p += 2)
        while (n % p == 0) { fac[nf++] = p; n /= p; }
    return nf;
}

static fft_plan *plan_create(int n)
{
    fft_plan *p = (fft_plan *)calloc(1, sizeof(fft_plan));
    p->n = n;
    p->nfac = factorize(n, p->fac);
    size_t ntw = 0;
    int Ns = 1;
    for (int s = 0; s < p->nfac; s++) { ntw += (size_t)(p->fac[s] - 1) * Ns; Ns *= p->fac[s]; }
    p->tw = (float *)malloc(sizeof(float) * 2 * (ntw + 1));
    for (int s = 0; s < p->nfac; s++) dft_table(p->fac[s]);
    Ns = 1;
    size_t o = 0;
    for (int s = 0; s < p->nfac; s++) {
        const int R = p->fac[s];
        for (int r = 1; r < R; r++)
            for (int k = 0; k < Ns; k++, o++) {
                const double a = -2.0 * M_PI * (double)r * k / ((double)Ns * R);
                p->tw[2 * o] = (float)cos(a);
                p->tw[2 * o + 1] = (float)sin(a);
            }
        Ns *= R;
    }
    return p;
}

static void plan_destroy(fft_plan *p)
{
    if (!p) return;
    free(p->tw);
    free(p);
}

/* cos/sin(2 pi e / R) for e < R, per radix (filled once; radices up to 64) */
static float g_dft_c[65][64], g_dft_s[65][64];
static int g_dft_ready[65];

static void dft_table(int R)
{
    if (R > 64 || g_dft_ready[R]) return;
#pragma omp critical(mdk_oracle_dft_table)
    if (!g_dft_ready[R]) {
        for (int e = 0; e < R; e++) {
            g_dft_c[R][e] = (float)cos(2.0 * M_PI * e / R);
            g_dft_s[R][e] = (float)sin(2.0 * M_PI * e / R);
        }
        g_dft_ready[R] = 1;
    }
}

/* small R-point DFT in place on v (interleaved), sign -1/+1 */
static void small_dft(int R, float *v, int sign)
{
    float out[2 * 64];
    if (R == 2) {
        float ar = v[0], ai = v[1], br = v[2], bi = v[3];
        v[0] = ar + br; v[1] = ai + bi; v[2] = ar - br; v[3] = ai - bi;
        return;
    }
    if (R == 4) {
        float a0r = v[0] + v[4], a0i = v[1] + v[5], a1r = v[0] - v[4], a1i = v[1] - v[5];
        float b0r = v[2] + v[6], b0i = v[3] + v[7], b1r = v[2] - v[6], b1i = v[3] - v[7];
        /* b1 * (sign * i) */
        float tr = -sign * b1i, ti = sign * b1r;
        v[0] = a0r + b0r; v[1] = a0i + b0i;
        v[4] = a0r - b0r; v[5] = a0i - b0i;
        v[2] = a1r + tr; v[3] = a1i + ti;
        v[6] = a1r - tr; v[7] = a1i - ti;
        return;
    }
    const float *ct = g_dft_c[R], *st = g_dft_s[R];
    for (int m = 0; m < R; m++) {
        float accr = v[0], acci = v[1];
        for (int r = 1, e = m; r < R; r++, e = (e + m) % R) {
            const float c = ct[e], s = sign * st[e];
            accr += v[2 * r] * c - v[2 * r + 1] * s;
            acci += v[2 * r] * s + v[2 * r + 1] * c;
        }
        out[2 * m] = accr;
        out[2 * m + 1] = acci;
    }
    memcpy(v, out, sizeof(float) * 2 * R);
}

/* one n-point transform; x and y are n-complex buffers; result returned in x */
static void fft_line(const fft_plan *p, float *x, float *y, int sign)
{
    const int n = p->n;
    int Ns = 1;
    const float *tw = p->tw;
    float v[2 * 64];
    for (int s = 0; s < p->nfac; s++) {
        const int R = p->fac[s], nr = n / R;
        for (int j = 0; j < nr; j++) {
            const int k = j % Ns;
            v[0] = x[2 * j];
            v[1] = x[2 * j + 1];
            for (int r = 1; r < R; r++) {
                const float xr = x[2 * (j + r * nr)], xi = x[2 * (j + r * nr) + 1];
                const float wr = tw[2 * ((r - 1) * Ns + k)];
                const float wi = -sign * tw[2 * ((r - 1) * Ns + k) + 1]; /* stored twiddles are the forward ones */
                v[2 * r] = xr * wr - xi * wi;
                v[2 * r + 1] = xr * wi + xi * wr;
            }
            small_dft(R, v, sign);
            const int d = (j / Ns) * Ns * R + k;
            for (int m = 0; m < R; m++) {
                y[2 * (d + m * Ns)] = v[2 * m];
                y[2 * (d + m * Ns) + 1] = v[2 * m + 1];
            }
        }
        tw += 2 * (size_t)(R - 1) * Ns;
        float *t = x; x = y; y = t;
        Ns *= R;
    }
    if (p->nfac & 1) memcpy(y, x, sizeof(float) * 2 * n); /* result must end in caller's x */
}

int or_fft1d(int n, int count, const float *in, float *out, int sign)
{
    fft_plan *p = plan_create(n);
    float *a = (float *)malloc(sizeof(float) * 4 * n), *b = a + 2 * n;
    for (int c = 0; c < count; c++) {
        memcpy(a, in + (size_t)2 * n * c, sizeof(float) * 2 * n);
        fft_line(p, a, b, sign);
        memcpy(out + (size_t)2 * n * c, a, sizeof(float) * 2 * n);
    }
    free(a);
    plan_destroy(p);
    return 0;
}

/* 3D transform on a caller-owned plan; along z (contiguous), y, then x.  The strided
 * axes are gathered FB lines at a time (FB consecutive z) for cache locality.
 * work: 4*n*(FB+1) floats. */
#define FB 16
static void fft3d_plan(const fft_plan *p, float *d, int sign, float *work)
{
    const int n = p->n;
    const size_t n2 = (size_t)n * n;
    float *a = work, *b = work + 2 * (size_t)n, *blk = work + 4 * (size_t)n; /* blk: FB lines */
    for (size_t l = 0; l < n2; l++) { /* z lines, in place */
        memcpy(a, d + 2 * l * n, sizeof(float) * 2 * n);
        fft_line(p, a, b, sign);
        memcpy(d + 2 * l * n, a, sizeof(float) * 2 * n);
    }
    for (int pass = 0; pass < 2; pass++) { /* pass 0: y lines, pass 1: x lines */
        const size_t outer = pass == 0 ? (size_t)n : 1, inner = pass == 0 ? (size_t)n : n2;
        const size_t stride = pass == 0 ? (size_t)n : n2; /* element stride along the axis */
        for (size_t o = 0; o < outer; o++)
            for (size_t z0 = 0; z0 < inner; z0 += FB) {
                const int w = (int)((inner - z0) < FB ? (inner - z0) : FB);
                float *base = d + 2 * (o * n2 + z0);
                for (int e = 0; e < n; e++)
                    for (int q = 0; q < w; q++) {
                        blk[2 * ((size_t)q * n + e)] = base[2 * (e * stride + q)];
                        blk[2 * ((size_t)q * n + e) + 1] = base[2 * (e * stride + q) + 1];
                    }
                for (int q = 0; q < w; q++) {
                    memcpy(a, blk + 2 * (size_t)q * n, sizeof(float) * 2 * n);
                    fft_line(p, a, b, sign);
                    memcpy(blk + 2 * (size_t)q * n, a, sizeof(float) * 2 * n);
                }
                for (int e = 0; e < n; e++)
                    for (int q = 0; q < w; q++) {
                        base[2 * (e * stride + q)] = blk[2 * ((size_t)q * n + e)];
                        base[2 * (e * stride + q) + 1] = blk[2 * ((size_t)q * n + e) + 1];
                    }
            }
    }
}

int or_fft3d(int nf, const float *in, float *out, int sign)
{
    const size_t nf3 = (size_t)nf * nf * nf;
    if (out != in) memcpy(out, in, sizeof(float) * 2 * nf3);
    fft_plan *p = plan_create(nf);
    float *work = (float *)malloc(sizeof(float) * 4 * (size_t)nf * (FB + 1));
    fft3d_plan(p, out, sign, work);
    free(work);
    plan_destroy(p);
    return 0;
}

/* ------------------------------------------------------------------ A9 */
void or_convolution(size_t nf3, const float *rr, const float *ri, const float *l, float *out)
{
    for (size_t i = 0; i < nf3; i++) {
        const float lr = l[2 * i], li = l[2 * i + 1];
        out[2 * i] = rr[i] * lr + ri[i] * li;
        out[2 * i + 1] = rr[i] * li - ri[i] * lr;
    }
}

/* ------------------------------------------------------------------ A11 */
void or_score_sort(int nf, int t, const float *fo, int rot, or_sortscore *sel)
{
    const int nf2 = nf * nf;
    const int nf3 = nf2 * nf;
    for (int i = 0; i < t; i++) {
        sel[i].score = -99999.0f;
        sel[i].index[1] = sel[i].index[2] = sel[i].index[3] = 0;
    }
    if (t != 1) { /* src/fft_process.cpp:415-431 */
        for (int i = 0; i < nf3; i++) {
            const float raw = fo[2 * (size_t)i] / nf3;
            if (raw < sel[t - 1].score) continue;
            for (int j = 0; j < t; j++)
                if (raw > sel[j].score) {
                    for (int k = t - 1; k > j; k--) sel[k] = sel[k - 1];
                    sel[j].score = raw;
                    sel[j].index[1] = i / nf2;
                    sel[j].index[2] = (i / nf) % nf;
                    sel[j].index[3] = i % nf;
                    break;
                }
        }
    } else { /* src/fft_process.cpp:433-446 */
        float best = 0.0f;
        int bi = 0;
        for (int i = 0; i < nf3; i++) {
            const float raw = fo[2 * (size_t)i];
            if (best < raw) { best = raw; bi = i; }
        }
        sel[0].score = best / nf3;
        sel[0].index[1] = bi / nf2;
        sel[0].index[2] = (bi / nf) % nf;
        sel[0].index[3] = bi % nf;
    }
    for (int i = 0; i < t; i++) sel[i].index[0] = rot;
}

int or_receptor_fft(int ng, int nf, const float *grid_r, const float *grid_i, float *rec_r,
                    float *rec_i)
{
    const size_t nf3 = (size_t)nf * nf * nf;
    float *c = (float *)malloc(sizeof(float) * 2 * nf3);
    or_embed(ng, nf, grid_r, grid_i, c);
    or_fft3d(nf, c, c, -1);
    for (size_t i = 0; i < nf3; i++) { rec_r[i] = c[2 * i]; rec_i[i] = c[2 * i + 1]; }
    free(c);
    return 0;
}

/* ------------------------------------------------------------------ A12 */
static int cmp_score(const void *pa, const void *pb)
{
    const or_sortscore *a = (const or_sortscore *)pa, *b = (const or_sortscore *)pb;
    if (a->score > b->score) return -1;
    if (a->score < b->score) return 1;
    for (int k = 0; k < 4; k++) {
        if (a->index[k] < b->index[k]) return -1;
        if (a->index[k] > b->index[k]) return 1;
    }
    return 0;
}

void or_sort_topn(or_sortscore *top, int count, int n)
{
    (void)n;
    qsort(top, (size_t)count, sizeof(or_sortscore), cmp_score);
}

/* ------------------------------------------------------------------ A13 */
int or_dock(int ng, int nf, float gw, int na, const float *coord, const float *center,
            const float *radius, const float *charge, float delta, const float *rec_r,
            const float *rec_i, const float *angles, const int *rot_ids, int n_rot, int t,
            int threads, or_sortscore *out)
{
    const size_t nf3 = (size_t)nf * nf * nf;
    const size_t ng3 = (size_t)ng * ng * ng;
    float *gc = (float *)malloc(sizeof(float) * ng);
    or_grid_coord(ng, gw, gc);
    fft_plan *p = plan_create(nf);
    if (threads < 1) threads = 1;
#pragma omp parallel num_threads(threads)
    {
        float *xyz = (float *)malloc(sizeof(float) * 3 * (size_t)na);
        float *gr = (float *)malloc(sizeof(float) * 2 * ng3), *gi = gr + ng3;
        float *fin = (float *)malloc(sizeof(float) * 2 * nf3);
        float *fout = (float *)malloc(sizeof(float) * 2 * nf3);
        float *work = (float *)malloc(sizeof(float) * 4 * (size_t)nf * (FB + 1));
#pragma omp for schedule(dynamic, 1)
        for (int q = 0; q < n_rot; q++) {
            float r[9];
            const int ang = rot_ids[q];
            or_rotation_coef(&angles[3 * ang], r);
            or_rotate_ligand(na, coord, center, r, xyz);
            or_ligand_voxel(ng, gw, na, xyz, radius, charge, gc, delta, gr, gi);
            or_embed(ng, nf, gr, gi, fin);
            fft3d_plan(p, fin, -1, work);                 /* forward, in place */
            or_convolution(nf3, rec_r, rec_i, fin, fout);
            fft3d_plan(p, fout, +1, work);                /* inverse */
            or_score_sort(nf, t, fout, ang, &out[(size_t)q * t]);
        }
        free(xyz); free(gr); free(fin); free(fout); free(work);
    }
    plan_destroy(p);
    free(gc);
    return 0;
}

/* A15 (-O): the real-space sums of Docking::output_detail for one pose, as the reference
 * forms them (src/docking.cpp:720-801): per axis s = (3 ng - t) mod 2 ng and the ligand
 * range [lo, hi) of :725-744; for l, m, n ascending the float products
 *   rpsc = rec_rpsc * lig_r,  elec = rec_i * lig_i,  ace = (rec_r - rec_rpsc) * lig_r
 * and each added, when negative or positive, to its float sum in that order (the decoys'
 * sums, :746-807), and each added to an unsplit float sum (the sums of the poses around the
 * input position, :646-654).  out = (rPSC+, rPSC-, ELEC+, ELEC-, ACE+, ACE-, rPSC, ELEC,
 * ACE); cnt (optional) the six counts of the split sums. */
void or_detail_sums_f32(int ng, const float *rec_r, const float *rec_i, const float *rec_rpsc, const float *lig_r,
                        const float *lig_i, const int trans[3], float out[9], int cnt[6])
{
    const int ng2 = ng * ng;
    int s[3], lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        s[a] = (3 * ng - trans[a]) % (2 * ng);
        if (s[a] < ng) {
            lo[a] = ng - s[a];
            hi[a] = ng;
        } else {
            lo[a] = 0;
            hi[a] = 2 * ng - s[a];
        }
    }
    float acc[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, all[3] = {0.0f, 0.0f, 0.0f};
    int n6[6] = {0, 0, 0, 0, 0, 0};
    for (int l = lo[0]; l < hi[0]; l++) {
        const int rx = ng2 * (l + s[0] - ng), lx = ng2 * l;
        for (int m = lo[1]; m < hi[1]; m++) {
            const int ry = rx + ng * (m + s[1] - ng), ly = lx + ng * m;
            for (int n = lo[2]; n < hi[2]; n++) {
                const int rz = ry + (s[2] + n - ng), lz = ly + n;
                const float p[3] = {rec_rpsc[rz] * lig_r[lz], rec_i[rz] * lig_i[lz],
                                    (rec_r[rz] - rec_rpsc[rz]) * lig_r[lz]};
                for (int k = 0; k < 3; k++) {
                    all[k] += p[k];
                    if (p[k] < 0) {
                        acc[2 * k + 1] += p[k];
                        n6[2 * k + 1]++;
                    } else if (p[k] > 0) {
                        acc[2 * k] += p[k];
                        n6[2 * k]++;
                    }
                }
            }
        }
    }
    for (int k = 0; k < 6; k++) {
        out[k] = acc[k];
        if (cnt) cnt[k] = n6[k];
    }
    for (int k = 0; k < 3; k++) out[6 + k] = all[k];
}
