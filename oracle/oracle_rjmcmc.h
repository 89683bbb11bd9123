/* Oracle RJMCMC interface (included from oracle.h).  TEST INFRASTRUCTURE ONLY. */
#ifndef BAMF_ORACLE_RJMCMC_H
#define BAMF_ORACLE_RJMCMC_H
#include <stdint.h>

/* Chain parameters; the fields of the MCMC struct read by make_Params
 * (source/c/cRJMCMC.cpp:14-50).  Same layout as bamf_mcmc_params. */
typedef struct {
    float psf_sigma; int32_t n_trials; int32_t n_burnin;
    float i_std_fg, i_std_bg, x_std_fg, x_std_bg;
    float p_burnin[8]; float p_trials[8];
    float rho; float bnd_out; int32_t grid_zoom; float split_std; float icutoff;
    float bg_std; float abbg_std; float dx; float drift_x; float drift_y; int32_t is_backg;
} orc_mcmc_params;

/* One record per jump call (burn-in and trials).  Same layout as bamf_step_record. */
typedef struct {
    int32_t kind;       /* dispatch: 0 move 1 split 2 merge 3 birth 4 death 5 gsplit 6 gmerge 7 backfore */
    int32_t jump_type;  /* reference JumpType of the record (1,2,3,10,4,5,6,7), 0 when none */
    int32_t flags;      /* bit0 accepted, bit1 chain record, bit2 grid update, bit3 bounds test failed */
    int32_t n;          /* N after the step */
    float llr, pr;      /* reference single-precision LLR = exp(sum) and PR */
    float bg;           /* BG after the step */
    float photons;      /* sum of I after the step (sequential single precision) */
    double log_pr;      /* log PR in double from the single-precision factors; NaN if undefined */
    int32_t draws;      /* uniforms consumed by the step */
    int32_t n_prop;     /* N of the proposed state */
    int32_t pix;        /* birth: residual-sampled pixel index (find_Birth's ii); else -1 */
    int32_t pad;
} orc_step;

enum { ORC_SRC_PHILOX = 0, ORC_SRC_TAPE = 1, ORC_SRC_LIBC = 2 };
typedef struct {
    int32_t kind; uint32_t roi; uint64_t seed;
    const uint32_t* tape; int64_t tape_len;   /* ORC_SRC_TAPE: raw 31-bit ints */
    uint32_t* record; int64_t record_cap;     /* optional log of the raw ints consumed */
    int64_t consumed; uint32_t step; uint32_t j; int32_t exhausted;
} orc_src;

typedef struct {
    orc_step* steps;                 /* n_burnin + n_trials entries, or NULL */
    float* pgrid; float* pback;      /* (ysize*zoom) x (xsize*zoom), or NULL */
    int32_t counters[30];            /* N_*, A_*, N_*Fg, A_*Fg, N_*Bg, A_*Bg  (jump, split, merge, birth, death) */
    int32_t n; float bg, abg, bbg;   /* final state */
    float* fin_i; float* fin_x; float* fin_y; int32_t* fin_sig; int32_t fin_cap;
    /* Chain / Chain_Test records (LibRJMCMC.cpp:290-322), or NULL */
    int64_t rec_cap, em_cap;
    int32_t* acc_head; float* acc_val; float* acc_em;
    int32_t* prop_head; float* prop_val; float* prop_em;
    int64_t n_rec, n_acc_em, n_prop_em;
    int32_t status;
    /* Thinned state snapshots (test infrastructure: reference posterior
     * samples).  After trials step t (0-based) with t + 1 > snap_skip and
     * (t + 1 - snap_skip) % snap_thin == 0, the current state is appended:
     * snap_n[k] = N, and N entries {X, Y, I} at snap_em (3 floats each).
     * snap_thin = 0: none. */
    int32_t snap_thin, snap_skip;
    int64_t snap_cap, snap_em_cap;
    int32_t* snap_n; float* snap_em;
    int64_t n_snap, n_snap_em;
} orc_out;

#define ORC_OK 0
#define ORC_E_UNSUPPORTED (-3)
#define ORC_E_NMAX (-4)
#define ORC_E_TAPE (-5)
#define ORC_E_ARG (-6)

int orc_rjmcmc(const float* data, int xsize, int ysize, const float* scmos,
               const orc_mcmc_params* p,
               const float* off_pdf, int off_len, const float* sig_pdf, int sig_len,
               const float* back_pdf, int back_len,
               int n0, const float* i0, const float* x0, const float* y0, float bg0,
               orc_src* src, orc_out* out);
/* The same chain with likelihood_ratio's model choice: opt_model 1 evaluates
 * NumModel on the sampled PSF psf of size psf_dims[4] (LibRJMCMC.cpp:62-170). */
int orc_rjmcmc_psf(const float* data, int xsize, int ysize, const float* scmos,
                   const orc_mcmc_params* p,
                   const float* off_pdf, int off_len, const float* sig_pdf, int sig_len,
                   const float* back_pdf, int back_len,
                   int n0, const float* i0, const float* x0, const float* y0, float bg0,
                   const float* psf, const int32_t* psf_dims, int opt_model,
                   orc_src* src, orc_out* out);

/* Pieces exposed for unit tests. */
float orc_urand_from_raw(uint32_t r);
float orc_stat_findpdf(const float* pdf, int len, float dx, float v);
int   orc_stat_tables(const float* pdf, int len, float dx, float* xarr, float* cdf,
                      float* uniq_x, float* uniq_cdf);   /* returns number of unique entries */
float orc_stat_genrand(const float* uniq_x, const float* uniq_cdf, int nuniq, float u);
void  orc_model(int n, const float* I, const float* X, const float* Y, float bg, float abg, float bbg,
                int xsize, int ysize, float sigma, const float* scmos, float* m);
void  orc_num_model(int n, const float* I, const float* X, const float* Y, float bg, float abg, float bbg,
                    int xsize, int ysize, const float* psf, const int32_t* dims, const float* scmos, float* m);
void  orc_gsplit_map(int np, const float* I, const float* X, const float* Y, const int* idx, float u1, float u2,
                     float u3, float* I1, float* X1, float* Y1, float* In, float* Xn, float* Yn, float* I0);
void  orc_gmerge_map(int np, const float* I, const float* X, const float* Y, int jj, const int* idx, float* I1,
                     float* X1, float* Y1, float* U1, float* U2, float* U3, float* I0);
void  orc_split_map(float I, float mux, float muy, float u1, float u2, float u3, float* out6);
void  orc_merge_map(float I1, float X1, float Y1, float I2, float X2, float Y2, float* out6);
/* Planted mistakes, for the oracle's pinning tests only (see rjmcmc.c). */
enum {
    ORC_SLIP_NONE = 0,
    ORC_SLIP_SPLIT_JACOB_POW1 = 1,   /* split: I / (1 - u1) in place of I / (1 - u1)^2 */
    ORC_SLIP_SPLIT_P_INVERTED = 2,   /* split: P[2] / P[1] */
    ORC_SLIP_SPLIT_PRM_N = 3,        /* split: MeanN / NEm in place of MeanN / (NEm + 1) */
    ORC_SLIP_SPLIT_NO_AREA = 4,      /* split: the 1 / Area position prior dropped */
    ORC_SLIP_MERGE_JACOB_POW3 = 5,   /* merge: Ijp / (1 - u1)^3 */
    ORC_SLIP_MERGE_PRM_N1 = 6,       /* merge: MeanN / (NEm + 1) in place of MeanN / NEm */
    ORC_SLIP_GSPLIT_JACOB_NP = 7,    /* gSplit: (1 - u3)^Np in place of ^(Np + 1) */
    ORC_SLIP_GSPLIT_US_NO_NP = 8,    /* gSplit: the 1 / Np density of u3 dropped */
    ORC_SLIP_GMERGE_P_INVERTED = 9,  /* gMerge: P[6] / P[5] */
    ORC_SLIP_BACKFORE_SWAPPED = 10   /* backFore: the two pdfs swapped */
};
void  orc_set_slip(int slip);
#endif
