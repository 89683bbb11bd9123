/*
 * Oracle: the per-ROI reversible-jump MCMC chain.  TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of class RJMCMC and its helpers in
 * source/c/LibRJMCMC.cpp and source/c/LibRJMCMCreturn.hpp, driven the way
 * mexFunction drives it (source/c/cRJMCMC.cpp:103-148).  Every uniform is
 * drawn in the reference's order (urand() LibRJMCMCreturn.hpp:17, randn()
 * :23-26) from a pluggable source: the counter-based Philox stream the GPU
 * chain also uses, a replay tape, or libc rand().
 *
 * Numeric semantics follow the reference as compiled by a C++11 compiler:
 *  - float overloads of erf/exp/log/sqrt/floor (<cmath> overloads);
 *  - pow(float, int) promotes to double (C++11 [c.math]), so the neighbour
 *    distance in closest()/closestMulti() is sqrt in double then cast to
 *    float, while distance() rounds the squared sum to float before sqrtf;
 *  - pi is the double literal 3.141592 (LibRJMCMCreturn.hpp:2);
 *  - no FMA contraction.
 * OptModel 1 (orc_rjmcmc_psf): NumModel (LibRJMCMC.cpp:62-170) in place of
 * Model, pinned by closed forms of the interpolation (tests/test_oracle_numpsf.py).
 * split/merge, gSplit/gMerge and backFore (LibRJMCMC.cpp:972-1255,
 * 1389-1627) are pinned by tests/test_oracle_split_merge.py and
 * tests/test_oracle_gsplit.py: numerical Jacobians of the proposal maps,
 * split-then-merge round trips whose log ratios cancel (the group pair to the
 * reference's (MeanN/(N+1))^2, its gMerge prior on N being inverted), split's
 * and merge's ratios assembled from independent pieces, the backFore label
 * law, and planted mistakes (orc_set_slip) that each of these checks catches.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define ORC_NMAX 512
#define PI_D 3.141592 /* LibRJMCMCreturn.hpp:2 */

enum { C_N = 0, C_A = 5, C_NFG = 10, C_AFG = 15, C_NBG = 20, C_ABG = 25 };
enum { J_MOVE = 0, J_SPLIT = 1, J_MERGE = 2, J_BIRTH = 3, J_DEATH = 4 };

typedef struct {
    int n;
    float bg, abg, bbg;
    float I[ORC_NMAX], X[ORC_NMAX], Y[ORC_NMAX];
    int sig[ORC_NMAX];
} em_t;

typedef struct {
    const float* pdf; int len; float* xarr; float* cdf; float* ux; float* uc; int nuniq;
} stat_t;

typedef struct {
    const float* data; const float* scmos; int X, Y;
    orc_mcmc_params p; float PVB[8], PVT[8], P[8];
    float meanN; int is_burnin;
    stat_t off, sig, back;
    em_t cur, T;
    float *M, *Mt;
    const float* psf; int32_t pdim[4]; int opt_model; /* SampledPSF, PSFsize, OptModel (cRJMCMC.cpp:35,73-74) */
    orc_src* src; orc_out* out; orc_step* st;
    int32_t draws;
    int err;
} ch_t;

/* ---------------- uniforms ---------------- */
float orc_urand_from_raw(uint32_t r) {
    /* (float)rand() / ((float)RAND_MAX + 1) with RAND_MAX = 2^31-1; (float)RAND_MAX + 1 == 2^31 */
    return (float)r / 2147483648.0f;
}

static uint32_t next_raw(ch_t* c) {
    orc_src* s = c->src;
    uint32_t r = 0;
    switch (s->kind) {
    case ORC_SRC_PHILOX: r = orc_chain_raw(s->seed, s->roi, s->step, s->j++); break;
    case ORC_SRC_TAPE:
        if (s->consumed < s->tape_len) r = s->tape[s->consumed];
        else { s->exhausted = 1; c->err = ORC_E_TAPE; r = 0; }
        break;
    default: r = (uint32_t)rand() & 0x7fffffffu; break;
    }
    if (s->record && s->consumed < s->record_cap) s->record[s->consumed] = r;
    s->consumed++;
    c->draws++;
    return r;
}
static float urand(ch_t* c) { return orc_urand_from_raw(next_raw(c)); }
/* LibRJMCMCreturn.hpp:23-26 (Irwin-Hall sum of 12 uniforms) */
static float randn(ch_t* c) {
    float tmp = 0;
    for (int nn = 0; nn < 12; nn++) tmp = tmp + urand(c);
    return tmp - 6.0f;
}
/* LibRJMCMCreturn.hpp:19-21 */
static float normpdf(float x, float u, float s) {
    float a = (float)sqrt(2.0f * PI_D);
    float e = expf(-(x - u) * (x - u) / (2 * s * s));
    return 1.0f / (a * s) * e;
}
/* LibRJMCMCreturn.hpp:28 */
static int kround(float a) { return (int)floor(a + .5); }
static int imin(int a, int b) { return a < b ? a : b; }

/* ---------------- Stat (LibRJMCMC.cpp:176-282) ---------------- */
int orc_stat_tables(const float* pdf, int len, float dx, float* xarr, float* cdf,
                    float* uniq_x, float* uniq_cdf) {
    float* xm = (float*)malloc(sizeof(float) * (len > 0 ? len : 1));
    cdf[0] = 0; xarr[0] = 0; xm[0] = 0;
    for (int ii = 1; ii < len; ii++) {
        cdf[ii] = cdf[ii - 1] + (pdf[ii] + pdf[ii - 1]) * dx / 2.0f;
        if (cdf[ii] > 1.0f) cdf[ii] = 1.0f;
        xm[ii] = cdf[ii];
        xarr[ii] = xarr[ii - 1] + dx;
    }
    for (int ii = len - 2; ii >= 0; ii--)
        if (cdf[ii] == cdf[ii + 1]) xm[ii + 1] = -1;
    int nu = 0;
    for (int ii = 0; ii < len; ii++)
        if (xm[ii] != -1) { uniq_x[nu] = xarr[ii]; uniq_cdf[nu] = cdf[ii]; nu++; }
    if (uniq_cdf[nu - 1] != 1) uniq_cdf[nu - 1] = 1;
    if (uniq_cdf[0] != 0) uniq_cdf[0] = 0;
    free(xm);
    return nu;
}

static float stat_findpdf_tab(const float* pdf, const float* xarr, int len, float I) {
    /* LibRJMCMC.cpp:219-246 */
    float IPDF;
    int Indd = -1;
    if (I >= xarr[len - 1] || I <= 0) return 0;
    for (int ji = 0; ji < len; ji++)
        if (xarr[ji] - I > 0) { Indd = ji; break; }
    if (Indd == -1 || Indd == 0 || Indd >= len) return 0;
    IPDF = (pdf[Indd] * (I - xarr[Indd - 1]) + pdf[Indd - 1] * (xarr[Indd] - I)) / (xarr[1] - xarr[0]);
    return IPDF;
}

float orc_stat_findpdf(const float* pdf, int len, float dx, float v) {
    float* xarr = (float*)malloc(sizeof(float) * len);
    xarr[0] = 0;
    for (int i = 1; i < len; i++) xarr[i] = xarr[i - 1] + dx;
    float r = stat_findpdf_tab(pdf, xarr, len, v);
    free(xarr);
    return r;
}

float orc_stat_genrand(const float* ux, const float* uc, int nu, float RandVal) {
    /* LibRJMCMC.cpp:248-282, with the uniform passed in */
    int SecInd = -1;
    if (RandVal <= 0) return 0;
    else if (RandVal >= 1) return ux[nu - 1];
    for (int ii = 0; ii < nu; ii++)
        if (uc[ii] - RandVal > 0) { SecInd = ii; break; }
    if (SecInd == -1) return 0;
    if (SecInd == 0) return 0;
    if (SecInd >= nu) return ux[nu - 1];
    return (ux[SecInd - 1] * (uc[SecInd] - RandVal) + ux[SecInd] * (RandVal - uc[SecInd - 1])) /
           (uc[SecInd] - uc[SecInd - 1]);
}

static void stat_init(stat_t* s, const float* pdf, int len, float dx) {
    s->pdf = pdf; s->len = len;
    s->xarr = (float*)malloc(sizeof(float) * len);
    s->cdf = (float*)malloc(sizeof(float) * len);
    s->ux = (float*)malloc(sizeof(float) * len);
    s->uc = (float*)malloc(sizeof(float) * len);
    s->nuniq = orc_stat_tables(pdf, len, dx, s->xarr, s->cdf, s->ux, s->uc);
}
static void stat_free(stat_t* s) { free(s->xarr); free(s->cdf); free(s->ux); free(s->uc); }
static float findpdf(const stat_t* s, float I) { return stat_findpdf_tab(s->pdf, s->xarr, s->len, I); }
static float genrand(ch_t* c, const stat_t* s) {
    float u = urand(c); /* drawn before the range checks (LibRJMCMC.cpp:251) */
    return orc_stat_genrand(s->ux, s->uc, s->nuniq, u);
}
static float pdf_of(ch_t* c, const em_t* e, int k) {
    return e->sig[k] == 1 ? findpdf(&c->sig, e->I[k]) : findpdf(&c->back, e->I[k]);
}

/* ---------------- model (LibRJMCMC.cpp:15-59) ---------------- */
void orc_model(int n, const float* I, const float* X, const float* Y, float bg, float abg, float bbg,
               int xsize, int ysize, float sigma, const float* scmos, float* m) {
    float Constant = 1 / (sigma * sqrtf(2.0f));
    float* nx = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1) * xsize);
    float* ny = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1) * ysize);
    for (int nn = 0; nn < n; nn++)
        for (int xx = 0; xx < xsize; xx++)
            nx[xsize * nn + xx] = I[nn] * 0.5f * (erff((xx - X[nn] + 1.0f) * Constant) - erff((xx - X[nn]) * Constant));
    for (int nn = 0; nn < n; nn++)
        for (int xx = 0; xx < ysize; xx++)
            ny[ysize * nn + xx] = 0.5f * (erff((xx - Y[nn] + 1.0f) * Constant) - erff((xx - Y[nn]) * Constant));
    for (int xx = 0; xx < xsize; xx++)
        for (int yy = 0; yy < ysize; yy++)
            m[yy + xx * ysize] = bg + abg * xx + bbg * yy + scmos[xx * ysize + yy];
    for (int nn = 0; nn < n; nn++)
        for (int yy = 0; yy < ysize; yy++)
            for (int xx = 0; xx < xsize; xx++)
                m[yy + xx * ysize] += nx[xsize * nn + xx] * ny[ysize * nn + yy];
    free(nx);
    free(ny);
}
/* ---------------- sampled-PSF model (NumModel, LibRJMCMC.cpp:67-169) ----------------
 * psf is the PSFsize(1) x PSFsize(2) x PSFsize(3) x PSFsize(4) array in
 * column-major order: entry (y, x, phase_y, phase_x) at y + P0 x + P01 phase_y
 * + P012 phase_x, the oversampling OS = PSFsize(4).  Each emitter is read off
 * the table at its sub-pixel phase and the next one (the phase after the last
 * being phase 0 of the next pixel), interpolated bilinearly with the ratios
 * formed in double; a linear index outside the table reads its last entry
 * (LibRJMCMC.cpp:131-158, where the reference also prints a message). */
static int num_clamp(int i, int n) { return (i >= n || i < 0) ? n - 1 : i; }
void orc_num_model(int n, const float* I, const float* X, const float* Y, float bg, float abg, float bbg,
                   int xsize, int ysize, const float* psf, const int32_t* dims, const float* scmos, float* m) {
    const double os = (double)dims[3];
    const int p0 = dims[0], p01 = dims[0] * dims[1], p012 = p01 * dims[2], p0123 = p012 * dims[3];
    for (int xx = 0; xx < xsize; xx++)
        for (int yy = 0; yy < ysize; yy++)
            m[yy + xx * ysize] = bg + abg * xx + bbg * yy + scmos[yy + xx * ysize];
    for (int nn = 0; nn < n; nn++) {
        const double xi = (double)X[nn], yi = (double)Y[nn];
        const int st0 = dims[0] / 2;
        const int stx = (int)ceil(st0 - xi), imx = (int)floor(os * (xi - floor(xi)));
        const int sty = (int)ceil(st0 - yi), imy = (int)floor(os * (yi - floor(yi)));
        int stx2, imx2, sty2, imy2;
        if (imx == os - 1) { stx2 = stx - 1; imx2 = 0; } else { stx2 = stx; imx2 = imx + 1; }
        if (imy == os - 1) { sty2 = sty - 1; imy2 = 0; } else { sty2 = sty; imy2 = imy + 1; }
        const int o1 = imy * p01 + imx * p012, o2 = imy * p01 + imx2 * p012;
        const int o3 = imy2 * p01 + imx * p012, o4 = imy2 * p01 + imx2 * p012;
        const float yr = (float)(1 - ((yi - floor(yi)) - (double)imy / os) * os);
        const float xr = (float)(1 - ((xi - floor(xi)) - (double)imx / os) * os);
        for (int xx = 0; xx < xsize; xx++)
            for (int yy = 0; yy < ysize; yy++) {
                const int i1 = num_clamp((sty + yy) + (stx + xx) * p0 + o1, p0123);
                const int i2 = num_clamp((sty + yy) + (stx2 + xx) * p0 + o2, p0123);
                const int i3 = num_clamp((sty2 + yy) + (stx + xx) * p0 + o3, p0123);
                const int i4 = num_clamp((sty2 + yy) + (stx2 + xx) * p0 + o4, p0123);
                const float A = xr * psf[i1] + (1.0f - xr) * psf[i2];
                const float B = xr * psf[i3] + (1.0f - xr) * psf[i4];
                m[yy + xx * ysize] = m[yy + xx * ysize] + I[nn] * (yr * A + (1.0f - yr) * B);
            }
    }
}
/* likelihood_ratio's choice of model (LibRJMCMC.cpp:1653-1671, 1695-1702) */
static void model_of(ch_t* c, const em_t* e, float* m) {
    if (c->opt_model == 1)
        orc_num_model(e->n, e->I, e->X, e->Y, e->bg, e->abg, e->bbg, c->X, c->Y, c->psf, c->pdim, c->scmos, m);
    else
        orc_model(e->n, e->I, e->X, e->Y, e->bg, e->abg, e->bbg, c->X, c->Y, c->p.psf_sigma, c->scmos, m);
}

/* LibRJMCMC.cpp:1650-1686.  Returns exp(sum); *sum_out = the sum. */
static float likelihood_ratio(ch_t* c, const em_t* t, float* sum_out) {
    model_of(c, t, c->Mt);
    int N = c->X * c->Y;
    float tmp = 0;
    for (int nn = 0; nn < N; nn++)
        tmp += c->M[nn] - c->Mt[nn] + c->data[nn] * logf(c->Mt[nn] / c->M[nn]);
    *sum_out = tmp;
    return expf(tmp);
}

/* ---------------- Emitters (LibRJMCMC.cpp:453-656) ---------------- */
static int em_add(ch_t* c, em_t* e, float I, float x, float y, int bf) {
    if (e->n >= ORC_NMAX) { c->err = ORC_E_NMAX; return 0; }
    e->I[e->n] = I; e->X[e->n] = x; e->Y[e->n] = y; e->sig[e->n] = bf; e->n++;
    return 1;
}
static void em_remove(em_t* e, int id) {
    if (e->n == 0) return;
    for (int k = id; k < e->n - 1; k++) {
        e->I[k] = e->I[k + 1]; e->X[k] = e->X[k + 1]; e->Y[k] = e->Y[k + 1]; e->sig[k] = e->sig[k + 1];
    }
    e->n--;
}
static int em_test(const em_t* e, int bnd, int xsize, int ysize) {
    for (int nn = 0; nn < e->n; nn++) if (e->I[nn] < 0) return 0;
    for (int nn = 0; nn < e->n; nn++) if (e->X[nn] < (-bnd)) return 0;
    for (int nn = 0; nn < e->n; nn++) if (e->Y[nn] < (-bnd)) return 0;
    for (int nn = 0; nn < e->n; nn++) if (e->X[nn] > (xsize + bnd)) return 0;
    for (int nn = 0; nn < e->n; nn++) if (e->Y[nn] > (ysize + bnd)) return 0;
    return 1;
}
/* sqrt(pow(dx,2)+pow(dy,2)) with pow(float,int) in double, stored to float */
static float nb_dist(const em_t* e, int nn, int jj) {
    double dx = (double)(e->X[nn] - e->X[jj]), dy = (double)(e->Y[nn] - e->Y[jj]);
    return (float)sqrt(dx * dx + dy * dy);
}
/* LibRJMCMC.cpp:525-558 */
static int em_closest(ch_t* c, const em_t* e, int jj, float dmin) {
    int ind[ORC_NMAX], num = 0;
    for (int nn = 0; nn < e->n; nn++)
        if (nn != jj && nb_dist(e, nn, jj) < dmin) ind[num++] = nn;
    if (num > 0) {
        float Ua = urand(c);
        int mm = imin((int)(Ua * num), num - 1);
        return ind[mm];
    }
    int k = 0;
    for (int nn = 0; nn < e->n; nn++) if (nn != jj) ind[k++] = nn;
    float Ua = urand(c);
    int mm = imin((int)(Ua * (e->n - 1)), e->n - 2);
    return ind[mm];
}
/* LibRJMCMC.cpp:615-650 */
static int em_closest_multi(ch_t* c, const em_t* e, int jj, float dmin, int* out) {
    int ind[ORC_NMAX], num = 0, np = 0;
    for (int nn = 0; nn < e->n; nn++)
        if (nn != jj && nb_dist(e, nn, jj) < dmin) ind[num++] = nn;
    if (num > 0) {
        float Ua = urand(c);
        if (Ua < 1.0f / (float)num) np = num;
        else np = imin((int)(Ua * num), num - 1);
        for (int nn = 0; nn < np; nn++) {
            Ua = urand(c);
            int k = imin((int)(Ua * (num - nn)), num - nn - 1);
            out[nn] = ind[k];
            for (int q = k; q < num - nn - 1; q++) ind[q] = ind[q + 1];
        }
    }
    return np;
}
/* LibRJMCMC.cpp:653-656 */
static float em_distance(const em_t* e, int jj, int ii) {
    double dx = (double)(e->X[ii] - e->X[jj]), dy = (double)(e->Y[ii] - e->Y[jj]);
    float d = (float)(dx * dx + dy * dy);
    return sqrtf(d);
}
static float nem(const em_t* e) {
    float s = 0;
    for (int nn = 0; nn < e->n; nn++) s = s + (float)e->sig[nn];
    return s;
}

/* ---------------- chain records (LibRJMCMC.cpp:290-322, 1629-1648) ---------------- */
static void add_rec(ch_t* c, const em_t* e, float llr, float pr, int type, int acc, int prop) {
    orc_out* o = c->out;
    if (!o->acc_head || o->n_rec >= o->rec_cap) return;
    int32_t* head = prop ? o->prop_head : o->acc_head;
    float* val = prop ? o->prop_val : o->acc_val;
    float* em = prop ? o->prop_em : o->acc_em;
    int64_t* nem_ = prop ? &o->n_prop_em : &o->n_acc_em;
    int64_t k = o->n_rec;
    head[k * 4 + 0] = e->n; head[k * 4 + 1] = type; head[k * 4 + 2] = acc; head[k * 4 + 3] = (int32_t)*nem_;
    val[k * 5 + 0] = llr; val[k * 5 + 1] = pr; val[k * 5 + 2] = e->bg; val[k * 5 + 3] = e->abg; val[k * 5 + 4] = e->bbg;
    if (e->n < 1) {
        if (*nem_ < o->em_cap) {
            float* q = em + *nem_ * 4;
            q[0] = -100; q[1] = -100; q[2] = 0; q[3] = -1;
        }
        (*nem_)++;
    } else {
        for (int nn = 0; nn < e->n; nn++) {
            if (*nem_ < o->em_cap) {
                float* q = em + *nem_ * 4;
                q[0] = e->X[nn] + c->p.drift_x; q[1] = e->Y[nn] + c->p.drift_y; q[2] = e->I[nn]; q[3] = (float)e->sig[nn];
            }
            (*nem_)++;
        }
    }
    if (prop) o->n_rec++;
}
static void record(ch_t* c, const em_t* t, float llr, float pr, int type, int acc) {
    add_rec(c, &c->cur, llr, pr, type, acc, 0);
    add_rec(c, t, llr, pr, type, acc, 1);
    c->st->jump_type = type;
    c->st->flags |= 2 | (acc ? 1 : 0);
    c->st->llr = llr;
    c->st->pr = pr;
    c->st->n_prop = t->n;
}
static void update_grids(ch_t* c) {
    c->st->flags |= 4;
    if (c->is_burnin || !c->out->pgrid) return;
    const em_t* e = &c->cur;
    int z = c->p.grid_zoom;
    for (int nn = 0; nn < e->n; nn++) {
        int x = kround((float)z * (e->X[nn] + c->p.drift_x));
        int y = kround((float)z * (e->Y[nn] + c->p.drift_y));
        float i = e->I[nn];
        if (x >= 0 && x < c->X * z && y >= 0 && y < c->Y * z && i > c->p.icutoff) {
            if (e->sig[nn] == 1) c->out->pgrid[(size_t)x * c->Y * z + y] += 1;
            else c->out->pback[(size_t)x * c->Y * z + y] += 1;
        }
    }
}
static void accept_state(ch_t* c) {
    c->cur = c->T;
    float* t = c->M; c->M = c->Mt; c->Mt = t;
}
static double lg(float v) { return log((double)v); }

/* Planted mistakes for the oracle's own pinning tests (0 = none): each one
 * changes one factor of one ratio, so a test can show that its check catches
 * that mistake.  Never set outside those tests. */
static int g_slip = 0;
void orc_set_slip(int slip) { g_slip = slip; }

/* split's proposal map (LibRJMCMC.cpp:1006-1021): emitter (I, mux, muy) with
 * (u1, u2, u3) -> out = {I1, X1, Y1} (the split emitter in place) and
 * {I2, X2, Y2} (the added one). */
void orc_split_map(float I, float mux, float muy, float u1, float u2, float u3, float* out) {
    out[0] = u1 * I;
    out[1] = mux + u2;
    out[2] = muy + u3;
    out[3] = (1 - u1) * I;
    out[4] = mux - u1 * u2 / (1 - u1);
    out[5] = muy - u1 * u3 / (1 - u1);
}

/* merge's map (LibRJMCMC.cpp:1141-1150): emitter jj (I1, X1, Y1) absorbs
 * ii (I2, X2, Y2) -> out = {I, mux, muy, u1, u2, u3}. */
void orc_merge_map(float I1, float X1, float Y1, float I2, float X2, float Y2, float* out) {
    float Ijp = I1 + I2;
    float muxjp = (I1 * X1 + I2 * X2) / Ijp;
    float muyjp = (I1 * Y1 + I2 * Y2) / Ijp;
    out[0] = Ijp;
    out[1] = muxjp;
    out[2] = muyjp;
    out[3] = I1 / Ijp;
    out[4] = X1 - muxjp;
    out[5] = Y1 - muyjp;
}

/* ---------------- jumps ---------------- */
/* LibRJMCMC.cpp:752-970 */
static void do_move(ch_t* c) {
    int JumpType = 1, Accept = 0;
    float LLR = 0, PR = 0, PrI = 1, PrBg = 1, IPDF = 1, IPDFtest = 1, sum = 0;
    int32_t* cnt = c->out->counters;
    em_t* E = &c->cur;
    em_t* T = &c->T;
    const orc_mcmc_params* p = &c->p;
    cnt[C_N + J_MOVE]++;
    *T = *E;
    if (E->n < 1) return;
    float Ua = urand(c);
    int nn = imin((int)(Ua * T->n), T->n - 1);
    float Ub = urand(c);
    if (Ub < 0.25) {
        JumpType = 2;
        T->bg = T->bg + p->bg_std * randn(c);
        T->abg = T->abg + p->abbg_std * randn(c);
        T->bbg = T->bbg + p->abbg_std * randn(c);
    } else if (Ub < 0.50 && Ub >= 0.25) {
        if (E->n < 2) return;
        JumpType = 10;
        int cl[ORC_NMAX];
        int np = em_closest_multi(c, E, nn, 2 * p->psf_sigma, cl);
        if (np == 0) return;
        float I0 = E->I[nn];
        for (int ii = 0; ii < np; ii++) I0 = I0 + E->I[cl[ii]];
        float II0 = 0, XI0 = 0, YI0 = 0, XII0 = 0, YII0 = 0;
        for (int ii = 0; ii < np; ii++) {
            int k = cl[ii];
            if (E->sig[k] == 1) {
                T->X[k] = E->X[k] + p->x_std_fg * randn(c);
                T->Y[k] = E->Y[k] + p->x_std_fg * randn(c);
                T->I[k] = (I0 / (float)(np + 1)) * (0.15f * randn(c) + 1);
                IPDFtest = findpdf(&c->sig, T->I[k]);
                IPDF = findpdf(&c->sig, E->I[k]);
            } else {
                T->X[k] = E->X[k] + p->x_std_bg * randn(c);
                T->Y[k] = E->Y[k] + p->x_std_bg * randn(c);
                T->I[k] = (I0 / (float)(np + 1)) * (0.15f * randn(c) + 1);
                IPDFtest = findpdf(&c->back, T->I[k]);
                IPDF = findpdf(&c->back, E->I[k]);
            }
            PrI = PrI * IPDFtest / IPDF;
            XI0 = XI0 + E->I[k] * E->X[k];
            YI0 = YI0 + E->I[k] * E->Y[k];
            XII0 = XII0 + T->I[k] * T->X[k];
            YII0 = YII0 + T->I[k] * T->Y[k];
            II0 = II0 + T->I[k];
        }
        T->I[nn] = I0 - II0;
        T->X[nn] = (XI0 + E->I[nn] * E->X[nn] - XII0) / T->I[nn];
        T->Y[nn] = (YI0 + E->I[nn] * E->Y[nn] - YII0) / T->I[nn];
        if (E->sig[nn] == 1) {
            IPDFtest = findpdf(&c->sig, T->I[nn]);
            IPDF = findpdf(&c->sig, E->I[nn]);
        } else {
            IPDFtest = findpdf(&c->back, T->I[nn]);
            IPDF = findpdf(&c->back, E->I[nn]);
        }
        PrI = PrI * IPDFtest / IPDF;
    } else if (Ub >= 0.5 && Ub < 0.75) {
        JumpType = 3;
        int ind[ORC_NMAX + 1];
        int NN = em_closest_multi(c, E, nn, 2 * p->psf_sigma, ind);
        ind[NN++] = nn;
        for (int ii = 0; ii < NN; ii++) {
            int k = ind[ii];
            if (E->sig[k] == 1) {
                cnt[C_NFG + J_MOVE]++;
                T->I[k] = T->I[k] + p->i_std_fg * randn(c);
                T->X[k] = T->X[k] + p->x_std_fg * randn(c);
                T->Y[k] = T->Y[k] + p->x_std_fg * randn(c);
                IPDFtest = findpdf(&c->sig, T->I[k]);
                IPDF = findpdf(&c->sig, E->I[k]);
            } else {
                cnt[C_NBG + J_MOVE]++;
                T->I[k] = T->I[k] + p->i_std_bg * randn(c);
                T->X[k] = T->X[k] + p->x_std_bg * randn(c);
                T->Y[k] = T->Y[k] + p->x_std_bg * randn(c);
                IPDFtest = findpdf(&c->back, T->I[k]);
                IPDF = findpdf(&c->back, E->I[k]);
            }
            PrI = PrI * IPDFtest / IPDF;
        }
    } else {
        if (E->sig[nn] == 1) {
            cnt[C_NFG + J_MOVE]++;
            T->I[nn] = T->I[nn] + p->i_std_fg * randn(c);
            T->X[nn] = T->X[nn] + p->x_std_fg * randn(c);
            T->Y[nn] = T->Y[nn] + p->x_std_fg * randn(c);
            IPDFtest = findpdf(&c->sig, T->I[nn]);
            IPDF = findpdf(&c->sig, E->I[nn]);
        } else {
            cnt[C_NBG + J_MOVE]++;
            T->I[nn] = T->I[nn] + p->i_std_bg * randn(c);
            T->X[nn] = T->X[nn] + p->x_std_bg * randn(c);
            T->Y[nn] = T->Y[nn] + p->x_std_bg * randn(c);
            IPDFtest = findpdf(&c->back, T->I[nn]);
            IPDF = findpdf(&c->back, E->I[nn]);
        }
        PrI = IPDFtest / IPDF;
    }
    if (!em_test(T, (int)p->bnd_out, c->X, c->Y)) {
        c->st->flags |= 8;
        record(c, T, LLR, PR, JumpType, Accept);
        update_grids(c);
        return;
    }
    LLR = likelihood_ratio(c, T, &sum);
    float BgPDF = findpdf(&c->off, E->bg);
    float BgPDFtest = findpdf(&c->off, T->bg);
    PrBg = (BgPDFtest / BgPDF) * expf(-(T->bbg * T->bbg - E->bbg * E->bbg) / 2) *
           expf(-(T->abg * T->abg - E->abg * E->abg) / 2);
    PR = LLR * PrI * PrBg;
    c->st->log_pr = (double)sum + lg(PrI) + lg(PrBg);
    Accept = (urand(c) < PR);
    if (Accept) {
        cnt[C_A + J_MOVE]++;
        if (E->sig[nn] == 1) cnt[C_AFG + J_MOVE]++; else cnt[C_ABG + J_MOVE]++;
        accept_state(c);
    }
    record(c, T, LLR, PR, JumpType, Accept);
    update_grids(c);
}

/* LibRJMCMC.cpp:972-1109 */
static void do_split(ch_t* c) {
    int JumpType = 4, Accept = 0;
    float LLR = 0, PR = 0, PrI = 1, PrX = 1, Us = 1, Jacob = 1, PrM = 0, sum = 0;
    int32_t* cnt = c->out->counters;
    em_t* E = &c->cur;
    em_t* T = &c->T;
    const orc_mcmc_params* p = &c->p;
    cnt[C_N + J_SPLIT]++;
    *T = *E;
    if (E->n < 1) return;
    float Ua = urand(c);
    int jj = imin((int)floorf((float)E->n * Ua), E->n - 1);
    float Ijp = E->I[jj], mux = E->X[jj], muy = E->Y[jj];
    int BF = E->sig[jj];
    if (BF == 1) cnt[C_NFG + J_SPLIT]++; else cnt[C_NBG + J_SPLIT]++;
    float u1 = urand(c);
    float u2 = p->split_std * randn(c);
    float u3 = p->split_std * randn(c);
    float sm[6];
    orc_split_map(Ijp, mux, muy, u1, u2, u3, sm);
    T->I[jj] = sm[0];
    T->X[jj] = sm[1];
    T->Y[jj] = sm[2];
    if (!em_add(c, T, sm[3], sm[4], sm[5], BF)) return;
    if (!em_test(T, (int)p->bnd_out, c->X, c->Y)) {
        c->st->flags |= 8;
        record(c, T, LLR, PR, JumpType, Accept);
        update_grids(c);
        return;
    }
    LLR = likelihood_ratio(c, T, &sum);
    float tmp1 = 1, tmp2 = 1;
    tmp1 = tmp1 * pdf_of(c, E, jj);
    tmp2 = tmp2 * pdf_of(c, T, jj);
    tmp2 = tmp2 * pdf_of(c, T, T->n - 1);
    float Norm1 = normpdf(u2, 0, p->split_std);
    float Norm2 = normpdf(u3, 0, p->split_std);
    float NEm = nem(E);
    PrM = c->meanN / (NEm + 1);
    PrI = tmp2 / tmp1;
    float Area = (float)((c->X + p->bnd_out) * (c->Y + p->bnd_out));
    PrX = 1 / Area;
    Us = 1 / (Norm1 * Norm2);
    double om = (double)(1 - u1);
    Jacob = (float)((double)Ijp / (om * om));
    float Pr12 = c->P[1] / c->P[2];
    if (g_slip == ORC_SLIP_SPLIT_JACOB_POW1) Jacob = (float)((double)Ijp / om);
    if (g_slip == ORC_SLIP_SPLIT_P_INVERTED) Pr12 = c->P[2] / c->P[1];
    if (g_slip == ORC_SLIP_SPLIT_PRM_N) PrM = c->meanN / NEm;
    if (g_slip == ORC_SLIP_SPLIT_NO_AREA) PrX = 1;
    PR = LLR * PrI * PrX * Us * Jacob * Pr12 * PrM;
    c->st->log_pr = (double)sum + lg(PrI) + lg(PrX) + lg(Us) + lg(Jacob) + lg(Pr12) + lg(PrM);
    Accept = (urand(c) < PR);
    if (Accept) {
        cnt[C_A + J_SPLIT]++;
        if (BF == 1) cnt[C_AFG + J_SPLIT]++; else cnt[C_ABG + J_SPLIT]++;
        accept_state(c);
    }
    record(c, T, LLR, PR, JumpType, Accept);
    update_grids(c);
}

/* LibRJMCMC.cpp:1111-1255 */
static void do_merge(ch_t* c) {
    int JumpType = 5, Accept = 0;
    float LLR = 0, PR = 0, PrI = 1, PrX = 1, Us = 1, Jacob = 1, PrM = 0, sum = 0;
    int32_t* cnt = c->out->counters;
    em_t* E = &c->cur;
    em_t* T = &c->T;
    const orc_mcmc_params* p = &c->p;
    cnt[C_N + J_MERGE]++;
    *T = *E;
    if (E->n < 2) { update_grids(c); return; }
    float Ua = urand(c);
    int jj = imin((int)floorf((float)E->n * Ua), E->n - 1);
    int ii = em_closest(c, E, jj, 3 * p->psf_sigma);
    if (E->sig[jj] == 1) cnt[C_NFG + J_MERGE]++; else cnt[C_NBG + J_MERGE]++;
    if (em_distance(E, ii, jj) > 3 * p->psf_sigma) { update_grids(c); return; }
    float mm[6];
    orc_merge_map(E->I[jj], E->X[jj], E->Y[jj], E->I[ii], E->X[ii], E->Y[ii], mm);
    float Ijp = mm[0], muxjp = mm[1], muyjp = mm[2], u1 = mm[3], u2 = mm[4], u3 = mm[5];
    T->I[jj] = Ijp;
    T->X[jj] = muxjp;
    T->Y[jj] = muyjp;
    float PDFfg = findpdf(&c->sig, Ijp);
    float PDFbg = findpdf(&c->back, Ijp);
    T->sig[jj] = 1;
    if (PDFbg > PDFfg) T->sig[jj] = 0;
    em_remove(T, ii);
    if (!em_test(T, (int)p->bnd_out, c->X, c->Y)) {
        c->st->flags |= 8;
        record(c, T, LLR, PR, JumpType, Accept);
        update_grids(c);
        return;
    }
    LLR = likelihood_ratio(c, T, &sum);
    float tmp1 = 1, tmp2 = 1;
    tmp1 = tmp1 * pdf_of(c, E, jj);
    tmp1 = tmp1 * pdf_of(c, E, ii);
    if (ii > jj) tmp2 = tmp2 * pdf_of(c, T, jj);
    else tmp2 = tmp2 * pdf_of(c, T, jj - 1);
    float Norm1 = normpdf(u2, 0, p->split_std);
    float Norm2 = normpdf(u3, 0, p->split_std);
    float NEm = nem(E);
    PrI = tmp1 / tmp2;
    float Area = (float)((c->X + p->bnd_out) * (c->Y + p->bnd_out));
    PrX = 1 / Area;
    Us = 1 / (Norm1 * Norm2);
    double om = (double)(1 - u1);
    Jacob = (float)((double)Ijp / (om * om));
    PrM = c->meanN / NEm;
    if (g_slip == ORC_SLIP_MERGE_JACOB_POW3) Jacob = (float)((double)Ijp / (om * om * om));
    if (g_slip == ORC_SLIP_MERGE_PRM_N1) PrM = c->meanN / (NEm + 1);
    PR = LLR / (PrI * PrX * Jacob * Us * PrM * c->P[1] / c->P[2]);
    c->st->log_pr = (double)sum - (lg(PrI) + lg(PrX) + lg(Jacob) + lg(Us) + lg(PrM) + lg(c->P[1]) - lg(c->P[2]));
    Accept = (urand(c) < PR);
    if (Accept) {
        if (E->sig[jj] == 1) cnt[C_AFG + J_MERGE]++; else cnt[C_ABG + J_MERGE]++;
        cnt[C_A + J_MERGE]++;
        accept_state(c);
    }
    record(c, T, LLR, PR, JumpType, Accept);
    update_grids(c);
}

/* LibRJMCMC.cpp:1688-1736 */
static void find_birth(ch_t* c, float* x, float* y, float* pixprob) {
    int N = c->X * c->Y, ii;
    float* tmp = c->Mt; /* scratch; the proposal model is rebuilt afterwards */
    for (int nn = 0; nn < N; nn++) {
        float v = c->data[nn] - c->M[nn];
        tmp[nn] = (0 > v) ? 0.0f : v;
    }
    for (int nn = 1; nn < N; nn++) tmp[nn] += tmp[nn - 1];
    int Ind = -1;
    float r = tmp[N - 1] * urand(c);
    for (ii = 0; ii < N; ii++) {
        Ind = Ind + 1;
        if (r <= tmp[ii]) break;
    }
    (void)Ind;
    c->st->pix = ii;
    pixprob[0] = 1;
    y[0] = (float)(ii - c->Y * (ii / c->Y)) + urand(c) - 0.5f;
    x[0] = (float)(int)floorf(((float)ii) / c->Y) + urand(c) - 0.5f;
}

/* LibRJMCMC.cpp:1258-1330 */
static void do_birth(ch_t* c) {
    int JumpType = 6, Accept = 0, BF;
    float LLR = 0, PR = 0, PrM, PixProb, x, y, sum = 0;
    int32_t* cnt = c->out->counters;
    em_t* E = &c->cur;
    em_t* T = &c->T;
    const orc_mcmc_params* p = &c->p;
    cnt[C_N + J_BIRTH]++;
    *T = *E;
    find_birth(c, &x, &y, &PixProb);
    if (p->is_backg == 1) BF = kround(urand(c));
    else BF = 1;
    float I_new;
    if (BF == 1) { cnt[C_NFG + J_BIRTH]++; I_new = genrand(c, &c->sig); }
    else { cnt[C_NBG + J_BIRTH]++; I_new = genrand(c, &c->back); }
    if (!em_add(c, T, I_new, x, y, BF)) return;
    if (!em_test(T, (int)p->bnd_out, c->X, c->Y)) {
        c->st->flags |= 8;
        record(c, T, LLR, PR, JumpType, Accept);
        update_grids(c);
        return;
    }
    LLR = likelihood_ratio(c, T, &sum);
    float NEm = nem(E);
    PrM = c->meanN / (NEm + 1);
    float xb = c->X + p->bnd_out, yb = c->Y + p->bnd_out;
    PR = LLR * c->P[3] / (c->P[4] * xb * yb * PixProb) * PrM;
    c->st->log_pr = (double)sum + lg(c->P[3]) - (lg(c->P[4]) + lg(xb) + lg(yb) + lg(PixProb)) + lg(PrM);
    Accept = (urand(c) < PR);
    if (Accept) {
        if (BF == 1) cnt[C_AFG + J_BIRTH]++; else cnt[C_ABG + J_BIRTH]++;
        cnt[C_A + J_BIRTH]++;
        accept_state(c);
    }
    record(c, T, LLR, PR, JumpType, Accept);
    update_grids(c);
}

/* LibRJMCMC.cpp:1332-1387 */
static void do_death(ch_t* c) {
    int JumpType = 7, Accept = 0;
    float LLR = 0, PR = 0, PrM, sum = 0;
    int32_t* cnt = c->out->counters;
    em_t* E = &c->cur;
    em_t* T = &c->T;
    cnt[C_N + J_DEATH]++;
    *T = *E;
    if (E->n < 1) return;
    float Ua = urand(c);
    int jj = imin((int)floorf((float)E->n * Ua), E->n - 1);
    if (E->sig[jj] == 1) cnt[C_NFG + J_DEATH]++; else cnt[C_NBG + J_DEATH]++;
    em_remove(T, jj);
    LLR = likelihood_ratio(c, T, &sum);
    float NEm = nem(E);
    PrM = NEm / c->meanN;
    float P43 = c->P[4] / c->P[3];
    PR = LLR * P43 * PrM;
    c->st->log_pr = (double)sum + lg(P43) + lg(PrM);
    Accept = (urand(c) < PR);
    if (Accept) {
        cnt[C_A + J_DEATH]++;
        if (E->sig[jj] == 1) cnt[C_AFG + J_DEATH]++; else cnt[C_ABG + J_DEATH]++;
        accept_state(c);
    }
    record(c, T, LLR, PR, JumpType, Accept);
    update_grids(c);
}

/* LibRJMCMC.cpp:1586-1627 */
static void do_backfore(ch_t* c) {
    em_t* E = &c->cur;
    if (E->n == 0) return;
    if (c->p.is_backg == 0) return;
    float Ua = urand(c);
    int nn = imin((int)floorf((float)E->n * Ua), E->n - 1);
    float BackP = findpdf(&c->back, E->I[nn]);
    float SignalP = findpdf(&c->sig, E->I[nn]);
    if (g_slip == ORC_SLIP_BACKFORE_SWAPPED) { float t = BackP; BackP = SignalP; SignalP = t; }
    if (E->sig[nn] == 1) {
        if (SignalP != 0) { if (BackP / SignalP > urand(c)) E->sig[nn] = 0; }
        else E->sig[nn] = 0;
    } else {
        if (BackP != 0) { if (SignalP / BackP > urand(c)) E->sig[nn] = 1; }
        else E->sig[nn] = 1;
    }
}

/* gSplit's proposal map (LibRJMCMC.cpp:1424-1431): the np group members
 * (idx[], the picked emitter last) move away from the new emitter, which
 * takes the fraction u3 of the group's intensity.  Writes the members' new
 * values to I1/X1/Y1[k] (k = 0..np-1, in idx order) and the new emitter to
 * *In, *Xn, *Yn.  I0/X0/Y0 are the group sums in idx order. */
void orc_gsplit_map(int np, const float* I, const float* X, const float* Y, const int* idx, float u1, float u2,
                    float u3, float* I1, float* X1, float* Y1, float* In, float* Xn, float* Yn, float* I0_out) {
    float I0 = 0, X0 = 0, Y0 = 0;
    for (int nn = 0; nn < np; nn++) {
        I0 = I0 + I[idx[nn]];
        X0 = X0 + X[idx[nn]];
        Y0 = Y0 + Y[idx[nn]];
    }
    for (int nn = 0; nn < np; nn++) {
        int k = idx[nn];
        X1[nn] = (X[k] - u3 * (X0 / (float)np + u1)) / (1 - u3);
        Y1[nn] = (Y[k] - u3 * (Y0 / (float)np + u2)) / (1 - u3);
        I1[nn] = (1 - u3) * I[k];
    }
    *In = u3 * I0;
    *Xn = X0 / (float)np + u1;
    *Yn = Y0 / (float)np + u2;
    *I0_out = I0;
}

/* gMerge's proposal map (LibRJMCMC.cpp:1509-1528): emitter jj is absorbed by
 * the np members idx[]; writes their new values to I1/X1/Y1[k] and returns
 * U1, U2, U3 and the group intensity I0 (jj first, then idx order). */
void orc_gmerge_map(int np, const float* I, const float* X, const float* Y, int jj, const int* idx, float* I1,
                    float* X1, float* Y1, float* U1, float* U2, float* U3, float* I0_out) {
    float I0 = I[jj];
    for (int nn = 0; nn < np; nn++) I0 = I0 + I[idx[nn]];
    float u3 = I[jj] / I0, XX0 = 0, YY0 = 0;
    for (int nn = 0; nn < np; nn++) {
        int k = idx[nn];
        I1[nn] = I[k] / (1 - u3);
        X1[nn] = u3 * X[jj] + (1 - u3) * X[k];
        XX0 = XX0 + u3 * X[jj] + (1 - u3) * X[k];
        Y1[nn] = u3 * Y[jj] + (1 - u3) * Y[k];
        YY0 = YY0 + u3 * Y[jj] + (1 - u3) * Y[k];
    }
    *U1 = X[jj] - XX0 / (float)np;
    *U2 = Y[jj] - YY0 / (float)np;
    *U3 = u3;
    *I0_out = I0;
}

/* LibRJMCMC.cpp:1389-1485.  No acceptance counters, no bounds test. */
static void do_gsplit(ch_t* c) {
    int JumpType = 8, Accept = 0;
    float LLR = 0, PR = 0, PrI = 1, PrX = 1, Us = 1, Jacob = 1, PrM = 0, sum = 0;
    em_t* E = &c->cur;
    em_t* T = &c->T;
    const orc_mcmc_params* p = &c->p;
    if (E->n < 2) return;
    *T = *E;
    float Ua = urand(c);
    int jj = imin((int)floorf((float)E->n * Ua), E->n - 1);
    int cl[ORC_NMAX + 1];
    int np = em_closest_multi(c, E, jj, 2 * p->psf_sigma, cl);
    cl[np++] = jj;
    if (np == 1) return;
    float u1 = (float)0.3 * randn(c);
    float u2 = (float)0.3 * randn(c);
    float u3 = (float)(1 / (float)np) * urand(c);
    float I1[ORC_NMAX + 1], X1[ORC_NMAX + 1], Y1[ORC_NMAX + 1], In, Xn, Yn, I0;
    orc_gsplit_map(np, E->I, E->X, E->Y, cl, u1, u2, u3, I1, X1, Y1, &In, &Xn, &Yn, &I0);
    for (int nn = 0; nn < np; nn++) { T->X[cl[nn]] = X1[nn]; T->Y[cl[nn]] = Y1[nn]; T->I[cl[nn]] = I1[nn]; }
    if (!em_add(c, T, In, Xn, Yn, 1)) return;
    LLR = likelihood_ratio(c, T, &sum);
    float tmpCurrent, tmpProposed;
    for (int nn = 0; nn < np; nn++) {
        tmpCurrent = pdf_of(c, E, cl[nn]);
        /* the proposed set's Signal is read at nn, not at Closest[nn] (LibRJMCMC.cpp:1447) */
        tmpProposed = T->sig[nn] == 1 ? findpdf(&c->sig, T->I[cl[nn]]) : findpdf(&c->back, T->I[cl[nn]]);
        PrI = PrI * tmpProposed / tmpCurrent;
    }
    tmpProposed = pdf_of(c, T, T->n - 1);
    PrI = PrI * (tmpProposed);
    float PDFu1 = normpdf(u1, 0.0f, 0.4f);
    float PDFu2 = normpdf(u2, 0.0f, 0.4f);
    float PDFu3 = (float)np;
    Us = 1 / (PDFu1 * PDFu2 * PDFu3);
    float NEm = nem(E);
    PrM = c->meanN / (NEm + 1);
    float Area = (float)((c->X + p->bnd_out) * (c->Y + p->bnd_out));
    PrX = 1 / Area;
    Jacob = I0 / powf(1 - u3, (float)(np + 1));
    float P56 = c->P[5] / c->P[6];
    if (g_slip == ORC_SLIP_GSPLIT_JACOB_NP) Jacob = I0 / powf(1 - u3, (float)np);
    if (g_slip == ORC_SLIP_GSPLIT_US_NO_NP) Us = 1 / (PDFu1 * PDFu2);
    PR = LLR * PrI * PrX * Us * Jacob * PrM * P56;
    c->st->log_pr = (double)sum + lg(PrI) + lg(PrX) + lg(Us) + lg(Jacob) + lg(PrM) + lg(P56);
    Accept = (urand(c) < PR);
    if (Accept) accept_state(c);
    record(c, T, LLR, PR, JumpType, Accept);
    update_grids(c);
}

/* LibRJMCMC.cpp:1487-1584.  No acceptance counters, no bounds test. */
static void do_gmerge(ch_t* c) {
    int JumpType = 9, Accept = 0;
    float LLR = 0, PR = 0, PrI = 1, PrX = 1, Us = 1, Jacob = 1, PrM = 0, sum = 0;
    em_t* E = &c->cur;
    em_t* T = &c->T;
    const orc_mcmc_params* p = &c->p;
    if (E->n < 3) return;
    *T = *E;
    float Ua = urand(c);
    int jj = imin((int)floorf((float)E->n * Ua), E->n - 1);
    int cl[ORC_NMAX];
    int np = em_closest_multi(c, E, jj, 3 * p->psf_sigma, cl);
    if (np == 0) return;
    float I1[ORC_NMAX], X1[ORC_NMAX], Y1[ORC_NMAX], U1, U2, U3, I0;
    orc_gmerge_map(np, E->I, E->X, E->Y, jj, cl, I1, X1, Y1, &U1, &U2, &U3, &I0);
    for (int nn = 0; nn < np; nn++) { T->I[cl[nn]] = I1[nn]; T->X[cl[nn]] = X1[nn]; T->Y[cl[nn]] = Y1[nn]; }
    float tmpProposed, tmpCurrent;
    for (int nn = 0; nn < np; nn++) {
        tmpProposed = pdf_of(c, T, cl[nn]);
        tmpCurrent = pdf_of(c, E, cl[nn]);
        PrI = PrI * (tmpCurrent / tmpProposed);
    }
    tmpCurrent = pdf_of(c, E, jj);
    PrI = PrI * tmpCurrent;
    em_remove(T, jj);
    LLR = likelihood_ratio(c, T, &sum);
    float PDFu1 = normpdf(U1, 0.0f, 0.4f);
    float PDFu2 = normpdf(U2, 0.0f, 0.4f);
    float PDFu3 = (float)np;
    Us = 1 / (PDFu1 * PDFu2 * PDFu3);
    float NEm = nem(E);
    PrM = NEm / c->meanN;
    float Area = (float)((c->X + p->bnd_out) * (c->Y + p->bnd_out));
    PrX = 1 / Area;
    Jacob = I0 / powf(1 - U3, (float)(np + 1));
    float P56 = c->P[5] / c->P[6];
    if (g_slip == ORC_SLIP_GMERGE_P_INVERTED) P56 = c->P[6] / c->P[5];
    PR = LLR / (PrI * PrX * Us * Jacob * PrM * P56);
    c->st->log_pr = (double)sum - (lg(PrI) + lg(PrX) + lg(Us) + lg(Jacob) + lg(PrM) + lg(P56));
    if (PDFu1 == 0 || PDFu2 == 0) {
        PR = -1;
        c->st->log_pr = -INFINITY;
    }
    Accept = (urand(c) < PR);
    if (Accept) accept_state(c);
    record(c, T, LLR, PR, JumpType, Accept);
    update_grids(c);
}

static void dispatch(ch_t* c, float rtmp, const float* PV, const float* PVB) {
    if (rtmp < PV[0]) { c->st->kind = 0; do_move(c); }
    else if (rtmp >= PV[0] && rtmp < PV[1]) { c->st->kind = 1; do_split(c); }
    else if (rtmp >= PV[1] && rtmp < PV[2]) { c->st->kind = 2; do_merge(c); }
    else if (rtmp >= PV[2] && rtmp < PV[3]) { c->st->kind = 3; do_birth(c); }
    else if (rtmp >= PV[3] && rtmp < PV[4]) { c->st->kind = 4; do_death(c); }
    else if (rtmp >= PVB[4] && rtmp < PVB[5]) { c->st->kind = 5; do_gsplit(c); }
    else if (rtmp >= PVB[5] && rtmp < PVB[6]) { c->st->kind = 6; do_gmerge(c); }
    else { c->st->kind = 7; do_backfore(c); }
}

int orc_rjmcmc(const float* data, int xsize, int ysize, const float* scmos,
               const orc_mcmc_params* p,
               const float* off_pdf, int off_len, const float* sig_pdf, int sig_len,
               const float* back_pdf, int back_len,
               int n0, const float* i0, const float* x0, const float* y0, float bg0,
               orc_src* src, orc_out* out) {
    return orc_rjmcmc_psf(data, xsize, ysize, scmos, p, off_pdf, off_len, sig_pdf, sig_len, back_pdf, back_len, n0,
                          i0, x0, y0, bg0, NULL, NULL, 0, src, out);
}

int orc_rjmcmc_psf(const float* data, int xsize, int ysize, const float* scmos,
                   const orc_mcmc_params* p,
                   const float* off_pdf, int off_len, const float* sig_pdf, int sig_len,
                   const float* back_pdf, int back_len,
                   int n0, const float* i0, const float* x0, const float* y0, float bg0,
                   const float* psf, const int32_t* psf_dims, int opt_model,
                   orc_src* src, orc_out* out) {
    if (opt_model == 1 && !(psf && psf_dims && psf_dims[0] > 0 && psf_dims[1] > 0 && psf_dims[2] > 0 &&
                            psf_dims[3] > 0))
        return ORC_E_ARG;
    ch_t* c = (ch_t*)calloc(1, sizeof(ch_t));
    c->opt_model = opt_model == 1 ? 1 : 0;
    c->psf = psf;
    if (c->opt_model) memcpy(c->pdim, psf_dims, sizeof(c->pdim));
    orc_step dummy;
    int P = xsize * ysize;
    c->data = data; c->scmos = scmos; c->X = xsize; c->Y = ysize; c->p = *p;
    c->src = src; c->out = out;
    /* make_Params cumulative vectors (cRJMCMC.cpp:37-47) */
    for (int ii = 0; ii < 8; ii++) {
        float b = 0, t = 0;
        for (int mm = 0; mm <= ii; mm++) { b = b + p->p_burnin[mm]; t = t + p->p_trials[mm]; }
        c->PVB[ii] = b; c->PVT[ii] = t;
    }
    c->meanN = p->rho * xsize * ysize; /* cRJMCMC.cpp:116 */
    stat_init(&c->off, off_pdf, off_len, p->dx);
    stat_init(&c->sig, sig_pdf, sig_len, p->dx);
    stat_init(&c->back, back_pdf, back_len, p->dx);
    c->M = (float*)malloc(sizeof(float) * P);
    c->Mt = (float*)malloc(sizeof(float) * P);
    memset(out->counters, 0, sizeof(out->counters));
    for (int k = 0; k < 5; k++) { out->counters[C_N + k] = 1; out->counters[C_NFG + k] = 1; out->counters[C_NBG + k] = 1; }
    out->n_rec = out->n_acc_em = out->n_prop_em = 0;
    out->n_snap = out->n_snap_em = 0;
    /* initial state (cRJMCMC.cpp:120-133) */
    c->cur.n = n0 > ORC_NMAX ? ORC_NMAX : n0;
    for (int k = 0; k < c->cur.n; k++) { c->cur.I[k] = i0[k]; c->cur.X[k] = x0[k]; c->cur.Y[k] = y0[k]; c->cur.sig[k] = 1; }
    c->cur.bg = bg0 == 0 ? 1 : bg0;
    c->cur.abg = 0; c->cur.bbg = 0;
    model_of(c, &c->cur, c->M);
    if (src->kind == ORC_SRC_LIBC) srand((unsigned)src->seed);
    src->consumed = 0; src->exhausted = 0;

    int total = p->n_burnin + p->n_trials;
    for (int s = 0; s < total && !c->err; s++) {
        int burn = s < p->n_burnin;
        c->is_burnin = burn;
        for (int k = 0; k < 8; k++) c->P[k] = burn ? p->p_burnin[k] : c->PVT[k];
        c->st = out->steps ? &out->steps[s] : &dummy;
        memset(c->st, 0, sizeof(orc_step));
        c->st->log_pr = NAN;
        c->st->pix = -1;
        c->draws = 0;
        src->step = (uint32_t)s;
        src->j = 0;
        float rtmp = urand(c);
        if (burn) dispatch(c, rtmp, c->PVB, c->PVB);
        else dispatch(c, rtmp, c->PVT, c->PVB);
        c->st->n = c->cur.n;
        c->st->bg = c->cur.bg;
        float ph = 0;
        for (int k = 0; k < c->cur.n; k++) ph = ph + c->cur.I[k];
        c->st->photons = ph;
        c->st->draws = c->draws;
        if (!(c->st->flags & 2)) c->st->n_prop = c->cur.n;
        if (!burn && out->snap_thin > 0) {
            int64_t t1 = (int64_t)(s - p->n_burnin) + 1 - out->snap_skip;
            if (t1 > 0 && t1 % out->snap_thin == 0 && out->n_snap < out->snap_cap) {
                out->snap_n[out->n_snap++] = c->cur.n;
                for (int k = 0; k < c->cur.n; k++) {
                    if (out->n_snap_em < out->snap_em_cap) {
                        float* q = out->snap_em + 3 * out->n_snap_em;
                        q[0] = c->cur.X[k]; q[1] = c->cur.Y[k]; q[2] = c->cur.I[k];
                    }
                    out->n_snap_em++;
                }
            }
        }
    }
    out->n = c->cur.n; out->bg = c->cur.bg; out->abg = c->cur.abg; out->bbg = c->cur.bbg;
    for (int k = 0; k < c->cur.n && k < out->fin_cap; k++) {
        if (out->fin_i) out->fin_i[k] = c->cur.I[k];
        if (out->fin_x) out->fin_x[k] = c->cur.X[k];
        if (out->fin_y) out->fin_y[k] = c->cur.Y[k];
        if (out->fin_sig) out->fin_sig[k] = c->cur.sig[k];
    }
    int err = c->err;
    out->status = err;
    stat_free(&c->off); stat_free(&c->sig); stat_free(&c->back);
    free(c->M); free(c->Mt);
    free(c);
    return err;
}
