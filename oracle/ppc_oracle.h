/*
 * ppc_oracle.h -- CPU restatement of the reference's photon-propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP kernel: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Nothing in ppc_amd/ links, loads or calls it.
 *
 * What it restates (each function cites the reference lines it follows):
 *   gpu/pro.cu:47-57    MWC draw                      orc_xrnd
 *   gpu/pro.cu:59-95    gamma / gaussian samplers     orc_mrnd, orc_grnd, orc_grnd2
 *   gpu/pro.cu:97-153   direction rotation            orc_rotate
 *   gpu/pro.cu:848-862  scattering angle              orc_scatter
 *   gpu/pro.cu:164-264  1-D / 2-D tilt                orc_zshift
 *   gpu/pro.cu:334-477  photon emission               orc_emit
 *   gpu/pro.cu:482-889  transport step loop           orc_transport
 *                       (with the GPU build's non-finite guard, pro.cu:727-729)
 *   gpu/ppc.cu:432-442  single-thread launch order    orc_run_xcpu
 *   (no counterpart)    score-function gradient       orc_adjoint_photon, orc_adjoint_path
 *
 * Pinning: every function but the two gradient ones is checked against the
 * reference's own CPU build (oracle/_ref, compiled from gpu/ppc.cu where it
 * lies) photon by photon, on random MWC states and bunches, and against closed
 * forms (tests/test_oracle*.py).  orc_adjoint_photon and orc_adjoint_path have no
 * counterpart in the reference; they are pinned by what the path density fixes
 * (tests/test_adjoint_oracle.py): the absorption and path-length scores by exact
 * identities (the per-layer lengths sum to the path length from the hit time;
 * straight paths give the per-layer moments), the scattering-event term (+la/(1-g)/sca
 * at each scattering) photon by photon (over a complete path the two scores of a layer
 * give its scatterings as whole numbers that add up to the counted ones) and by the
 * Poisson law of the scatterings (per layer E[n - sca l] = 0; in homogeneous ice
 * E[L] = 1/a and Var[n - s L] = s/a).
 * The tables it consumes (orc_tables) are filled either from the reference
 * build (ref_export) or from the product's table builder; both are compared
 * field by field in the tests.
 */
#ifndef PPC_ORACLE_H
#define PPC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* capacities of the reference (gpu/ini.cxx:1-26) */
#define ORC_LMAX 80
#define ORC_LYRS 172
#define ORC_CTX 11
#define ORC_CTY 10
#define ORC_CX 21
#define ORC_CY 19
#define ORC_NSTR 240
#define ORC_WNUM 512
#define ORC_MAXLYS 172
#define ORC_MAXGEO 16384

typedef struct {
  float R, F;
  float x, y, z;
} orc_dom;

typedef struct {
  int n, max;
  float x, y, r;
  float h, d;
  float dl, dh;
} orc_line;

typedef struct {
  /* layers */
  int size;
  float dh, rdh, hmin;
  /* scattering / DOM / hole ice */
  float ocv, sf, g, gr, xR;
  float SF, G, GR, hr, hr2, hs, ha;
  float hifl;
  /* anisotropy + birefringence */
  float azx, azy, k1, k2, kz, fr;
  float sum, bfr[12];
  float az_k1[ORC_MAXLYS], az_k2[ORC_MAXLYS], az_ra[ORC_MAXLYS], az_rb[ORC_MAXLYS];
  /* cell grid */
  float rx;
  int cn[2];
  float cl[2], crst[2], cb[4]; /* cb[i*2+j] = cb[i][j] */
  uint16_t ls[ORC_NSTR];
  uint16_t is_[ORC_CX * ORC_CY]; /* is[i][j] at i*ORC_CY+j */
  int nsc;
  orc_line sc[ORC_NSTR];
  /* tilt */
  int tmod, vthk, lnum, lpts, l0;
  float lmin, lrdz, r0, lnx, lny;
  float lr[ORC_LMAX];
  float mmin[2], mstp[2];
  int mnum[2], mcut[2];
  int mcol[ORC_CTX * ORC_CTY]; /* mcol[i][j] at i*ORC_CTY+j */
  float lp_x[ORC_LMAX * ORC_LYRS], lp_y[ORC_LMAX * ORC_LYRS];
  /* wavelength tables */
  int wvl[ORC_WNUM];
  float ocm[ORC_WNUM], coschr[ORC_WNUM], sinchr[ORC_WNUM];
  float abs_[ORC_WNUM * ORC_MAXLYS], sca[ORC_WNUM * ORC_MAXLYS];
  float wvs_w[ORC_WNUM], wvs_i[ORC_WNUM], wvs_f[ORC_WNUM];
  float eff;
  /* DOMs */
  int gsize;
  orc_dom oms[ORC_MAXGEO];
  int str[ORC_MAXGEO], dom[ORC_MAXGEO];
  /* RNG multipliers */
  int rsize;
} orc_tables;

/* one photon bunch (gpu/ini.cxx:399-419), flattened: the flasher fields (ka..ofla)
 * are used when type > 0, the particle fields (a..tau) otherwise. */
typedef struct {
  float r[4];
  float n[4];
  uint32_t q, num;
  int32_t type;
  float f;
  float a, b, beta, tau;
  float ka, up, fldr;
  int32_t fla, ofla;
} orc_bunch;

/* emitted photon (gpu/ini.cxx:391-397) */
typedef struct {
  float r[4];
  float n[4];
  uint32_t q, i;
  int32_t fla, ofla;
} orc_pbuf;

/* raw hit (gpu/ini.cxx:367-371) */
typedef struct {
  uint32_t i, n, z;
  float t;
  float pth, pph, dth, dph;
} orc_hit;

typedef struct {
  uint32_t x, c, a;
} orc_rng;

typedef struct {
  uint64_t photons, steps, crossings, cells, strings, doms, scatters, hits, draws, killed;
  uint64_t nonfinite; /* photons ended by the guard against non-finite optical depths */
} orc_counters;

/* gradient of the LLH weight of a photon's hits by per-layer absorption and
 * scattering parameters (score function of the path density; not part of the
 * reference).  The caller fills the inputs; ga, gb [size] accumulate. */
typedef struct {
  const float* lk;     /* [WNUM] d(abs)/d(p) of a wavelength bin (its lk factor) */
  const float* la;     /* [WNUM] d(sca)/d(s1) of a wavelength bin (its la factor) */
  const float* dpda;   /* [size] d(p)/d(a) */
  const float* ds1db;  /* [size] d(s1)/d(b) */
  const float* weight; /* [nseg][gsize] LLH weight of a hit of a DOM by segment */
  uint32_t nseg, gsize;
  double* ga;          /* [size] += d/d(a) */
  double* gb;          /* [size] += d/d(b) */
  double G;            /* working: weight of the hits still to come */
} orc_adjoint;

/* table handling (the struct is ~1.5 MB: allocate it on the heap) */
orc_tables* orc_tables_new(void);
void orc_tables_free(orc_tables* t);
/* pointer and size in bytes of a named field (NULL if unknown) */
void* orc_field(orc_tables* t, const char* name, uint64_t* bytes);

float orc_xrnd(orc_rng* s);
float orc_mrnd(float k, orc_rng* s);
float orc_grnd(orc_rng* s);
void orc_rotate(float cs, float si, float n[4], orc_rng* s);
float orc_scatter(float sf, float g, float gr, orc_rng* s);
float orc_zshift(const orc_tables* t, const float r[4], float* dz);

void orc_emit(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t idx, orc_rng* s, orc_pbuf* out);
/* returns the number of hits produced (written up to cap) */
int orc_transport(const orc_tables* t, const orc_pbuf* f, orc_rng* s, orc_hit* hits, int cap,
                  orc_counters* cnt);
/* one emitted+transported photon from a given state (the replay unit) */
int orc_photon(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t idx, orc_rng* s,
               orc_hit* hits, int cap, orc_counters* cnt);
/* the photon from state s0 run twice: once for its hits (their summed weight),
 * once accumulating the gradient; returns the number of hits */
int orc_adjoint_photon(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t idx, const orc_rng* s0,
                       orc_adjoint* A);
/* the photon from state s0 with its complete path scored (weight 1 from emission to
 * absorption, whatever it hits); row receives its wavelength row; returns its hits */
int orc_adjoint_path(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t idx, const orc_rng* s0,
                     orc_adjoint* A, orc_counters* cnt, uint32_t* row);
/* the reference's single-thread launch: nthr threads share one MWC stream and
 * thread k emits then transports photons k, k+nthr, ... (gpu/ppc.cu:439-440) */
uint32_t orc_run_xcpu(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t num, uint32_t nthr,
                      orc_rng* s, orc_hit* hits, uint32_t cap, orc_counters* cnt);

#ifdef __cplusplus
}
#endif

#endif
