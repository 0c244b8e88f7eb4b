/*
 * ppc_hip.h -- C-ABI of the MI355X photon-propagation library (libppc_hip.so).
 *
 * This is the seam the reference's CUDA backend fills in gpu/ppc.cu: the host
 * side (tables, front-ends, hit post-processing) hands packed tables and photon
 * bunches to the device and collects raw hits.  Plain C types only; every
 * function returns 0 on success or a negative ppc_hip_status code, with a
 * message available from ppc_hip_last_error().
 *
 * One context drives one GPU.  A context is not thread safe; calls on one
 * context must come from one host thread (as in the reference, gpu/ppc.cu).
 */
#ifndef PPC_HIP_H
#define PPC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ppc_hip_status {
  PPC_HIP_OK = 0,
  PPC_HIP_E_ARG = -1,      /* invalid argument */
  PPC_HIP_E_DEVICE = -2,   /* HIP runtime error (the reference exits with code 2) */
  PPC_HIP_E_NOMEM = -3,    /* device memory exhausted */
  PPC_HIP_E_STATE = -4,    /* call out of order (e.g. launch before upload) */
  PPC_HIP_E_CAPACITY = -5, /* a size exceeds the configured capacity */
  PPC_HIP_E_INCOMPLETE = -6 /* a launch ended with waves that did not finish or photons not
                               started (the reference exits with code 4, gpu/ppc.cu:266) */
};

/* photon bunch: the reference's `photon` (gpu/ini.cxx:399-419), 64 bytes */
typedef struct {
  float r[4];     /* start position [m], time [ns] */
  float n[4];     /* direction; n[3] = track length (particles), 0/1 for flashers */
  uint32_t q;     /* track segment number, copied to hits */
  uint32_t num;   /* number of photons in this bunch */
  int32_t type;   /* >0: flasher/LED (1..4); <=0: particle (-type of the light yield model) */
  float f;        /* fraction of light from the bare muon */
  union {
    struct {
      float a, b;   /* longitudinal profile: offset = b * Gamma(a) */
      float beta;   /* particle velocity */
      float tau;    /* luminescence decay time */
    } p;
    struct {
      float ka, up; /* beam width parameter, LED elevation [rad] */
      float fldr;   /* LED #1 azimuth [deg] (<0: azimuthally symmetric) */
      int16_t fla, ofla; /* flashing DOM index, DOM excluded from hits */
    } l;
  } u;
} ppc_hip_bunch;

/* raw hit: the reference's `hit` (gpu/ini.cxx:367-371), 32 bytes */
typedef struct {
  uint32_t i;     /* DOM index */
  uint32_t n;     /* track segment number (bunch q) */
  uint32_t z;     /* wavelength bin */
  float t;        /* arrival time [ns] */
  float pth, pph; /* photon direction at arrival (zenith, azimuth) */
  float dth, dph; /* impact point on the DOM (zenith, azimuth) */
} ppc_hip_hit;

/* string (line) of the cell grid: the reference's `line` (gpu/ini.cxx:438-443) */
typedef struct {
  int32_t n, max; /* first DOM index, number of DOMs - 1 */
  float x, y, r;  /* centre and search radius */
  float h, d;     /* highest DOM z, 1/spacing */
  float dl, dh;   /* low/high z spread around the regular spacing */
} ppc_hip_line;

/* DOM: the reference's `DOM` (gpu/ini.cxx:42-45) */
typedef struct {
  float R, F;  /* radius, vertical/horizontal axis ratio (<0: cylinder) */
  float r[3];  /* position [m] */
} ppc_hip_dom;

#define PPC_HIP_WNUM 512
#define PPC_HIP_MAXLYS 172
#define PPC_HIP_LMAX 80
#define PPC_HIP_LYRS 172
#define PPC_HIP_NSTR 240
#define PPC_HIP_CX 21
#define PPC_HIP_CY 19
#define PPC_HIP_CTX 11
#define PPC_HIP_CTY 10

/*
 * Tables built by the host from the ice directory (what gpu/ini.cxx:578-1700
 * puts into `dats`/`datz`), with the per-wavelength ice tables factorised:
 *   sca(bin, layer) = lay[layer].s0 + lay[layer].s1 * wv[bin].la
 *   abs(bin, layer) = lay[layer].p  * wv[bin].lk + lay[layer].t * wv[bin].abl
 * which restates gpu/ini.cxx:1677-1689 term by term.
 */
typedef struct {
  float s0, s1, p, t;
} ppc_hip_layer;

typedef struct {
  float k1, k2, ra, rb; /* gpu/ini.cxx:431-436 */
} ppc_hip_aniz;

typedef struct {
  float ocm;            /* group index / c [ns/m] */
  float coschr, sinchr; /* Cherenkov angle */
  float la, lk, abl;    /* wavelength factors of the factorised ice tables */
  int32_t wvl;          /* bin number written into hits */
  float pad;
} ppc_hip_wv;

typedef struct {
  /* layers (gpu/ini.cxx:1543-1562) */
  int32_t size;
  float dh, rdh, hmin;
  /* light / scattering / DOMs (gpu/ini.cxx:646-691) */
  float ocv, sf, g, gr, xR;
  float SF, G, GR, hr, hr2, hs, ha, hifl;
  float azx, azy, k1, k2, kz, fr;
  float sum, bfr[12];
  /* cell grid (gpu/ini.cxx:1083-1174) */
  float rx;
  int32_t cn[2];
  float cl[2], crst[2], cb[4]; /* cb[i*2+j] */
  /* tilt (gpu/ini.cxx:1178-1319) */
  int32_t tmod, vthk, lnum, lpts;
  float lmin, lrdz, r0, lnx, lny;
  float lr[PPC_HIP_LMAX];
  float mmin[2], mstp[2];
  int32_t mnum[2], mcut[2];
  int32_t mcol[PPC_HIP_CTX * PPC_HIP_CTY];
  /* arrays (host pointers; copied at upload) */
  const ppc_hip_layer* lay;  /* [size] */
  const ppc_hip_aniz* az;    /* [size] */
  const ppc_hip_wv* wv;      /* [PPC_HIP_WNUM] */
  const uint16_t* is;        /* [PPC_HIP_CX*PPC_HIP_CY], is[i*CY+j] */
  const uint16_t* ls;        /* [PPC_HIP_NSTR] */
  int32_t nsc;
  const ppc_hip_line* sc;    /* [nsc] */
  const float* lp;           /* [lnum][PPC_HIP_LYRS][2] tilt table (value, slope) */
  int32_t gsize;
  const ppc_hip_dom* oms;    /* [gsize] */
  int32_t rsize;
  const uint32_t* rm;        /* [rsize] MWC multipliers, one per concurrent lane */
} ppc_hip_tables;

/* launch shape and capacities (the reference derives these from the device,
 * gpu/ppc.cu:141-186; 0 selects the library default) */
typedef struct {
  int32_t blocks;          /* workgroups; default: 1792/threads per CU (28 waves per CU, 7 per SIMD) */
  int32_t threads;         /* lanes per workgroup (multiple of 64); default 256 */
  int32_t refill;          /* refill a wave once this many of its 64 lanes are idle; default 10 */
  int32_t deterministic;   /* 1: photon i always runs on lane i % lanes (the reference's DTMN) */
  uint32_t max_bunches;    /* bunch buffer capacity; default 1<<20 */
  uint32_t max_hits;       /* hit buffer capacity; default 1<<24 */
  uint32_t rec_stride;     /* record the MWC state at emission of every rec_stride-th photon (0: off) */
  uint32_t rec_max;        /* capacity of the record buffer */
  int32_t count_events;    /* 1: also count layer crossings, string and DOM tests, scatterings (slower) */
  int32_t record_starts;   /* 1: record the start state of every photon that hits (adjoint), max_hits of them */
  uint32_t chunk;          /* photons a wave claims per atomic on the work counter; 0: num/(128 waves), 64..1024 */
  int32_t tilt_lds;        /* 2-d tilt table: 0 in LDS when it fits in 16 KB, 1 in LDS when it fits, -1 in HBM (the 1-d tilt table is always read from HBM) */
  int32_t no_clearance;    /* 1: search the cell grid at every step (no skipping of steps far from all strings) */
} ppc_hip_config;

/* per-launch results (what gpu/ppc.cu:255-283 reads back) */
typedef struct {
  uint64_t photons;   /* photons propagated */
  uint64_t steps;     /* transport steps (loop iterations) */
  uint64_t hits;      /* raw hits produced (may exceed max_hits: overflow) */
  uint64_t stored;    /* hits stored in the buffer */
  uint64_t killed;    /* photons killed at emission (ofla < -1) */
  uint64_t nonfinite; /* photons dropped because TOT/SCA became inf/nan */
  float device_ms;    /* kernel time between HIP events on the launch stream */
  int32_t lanes;      /* lanes of the launch */
  /* event counts, filled when count_events is set */
  uint64_t crossings; /* layer boundaries crossed in the layer walk */
  uint64_t strings;   /* string (cylinder) intersection tests */
  uint64_t doms;      /* DOM (ellipsoid) intersection tests */
  uint64_t scatters;  /* scatterings */
  uint64_t hit_photons; /* photons that hit (record_starts; records beyond max_hits are dropped) */
  /* in-kernel time from the launch's first wave start to its first and to its last wave end,
   * by the device's wall clock (the reference's in-kernel span, gpu/ppc.cu:267-273,309) */
  float span_min_ms, span_max_ms;
} ppc_hip_stats;

/* replay record: MWC state of a lane when it starts emitting photon `photon` */
typedef struct {
  uint32_t photon;
  uint32_t x, c, a;
} ppc_hip_record;

/* a photon that hit (record_starts), to be replayed by the adjoint: its index in
 * the launch and the MWC state and lane it started from */
typedef struct {
  uint32_t photon;
  uint32_t x0, c0, lane;
} ppc_hip_adj_record;

typedef struct ppc_hip_ctx ppc_hip_ctx;

/* number of visible devices (listDevices, gpu/ppc.cu:404-425) */
int ppc_hip_device_count(int32_t* count);
/* name, CU count and memory of a device */
int ppc_hip_device_info(int32_t device, char* name, size_t name_len, int32_t* cus, uint64_t* mem);

/* choose(device) + gpu::gpu (gpu/ppc.cu:117-149, 490-508) */
int ppc_hip_open(int32_t device, const ppc_hip_config* cfg, ppc_hip_ctx** ctx);
/* free all device memory (gpu::fin + stop, gpu/ppc.cu:239-249,308-311) */
int ppc_hip_close(ppc_hip_ctx* ctx);

/* lanes of a launch; the tables must carry at least this many multipliers */
int ppc_hip_lanes(const ppc_hip_ctx* ctx, int32_t* lanes);
/* maximum bunches and hits per launch (getMaxBunchSize, gpu/ppc.cu:368) */
int ppc_hip_capacity(const ppc_hip_ctx* ctx, uint32_t* max_bunches, uint32_t* max_hits);

/* the clearance map the kernel uses to skip the DOM search on steps far from every string
 * (no counterpart in the reference; host only): squares of side `step` over the strings'
 * footprint + 100 m, origin (x0, y0), nx * ny entries (map == NULL: sizes only), each a
 * lower bound of the horizontal distance from any point of the square to the nearest string
 * cylinder (line radius, or hole-ice radius if larger), capped at 100 m; ppc_hip_upload
 * builds it with squares of 1 m (2 m, 4 m, ... for footprints too large for 4M entries) */
int ppc_hip_clearance_map(const ppc_hip_tables* tables, float step, float* map, size_t cap, int32_t* nx, int32_t* ny,
                          float* x0, float* y0);
/* copy tables to the device (gpu::ini, gpu/ppc.cu:151-237) */
int ppc_hip_upload(ppc_hip_ctx* ctx, const ppc_hip_tables* tables);
/* reset every lane's MWC state to x = 362436069 + sv, c = 1234567 (rs_ini, gpu/ini.cxx:556-571) */
int ppc_hip_seed(ppc_hip_ctx* ctx, uint32_t sv);
/* seed != 0: in the launches that follow, photon i starts from an MWC state of its own, a
 * hash of (seed, i), on the multiplier of the lane that runs it (lane i % lanes in
 * deterministic launches), instead of continuing its lane's stream; launches with the same
 * seed then use the same random numbers photon by photon (common random numbers for
 * differences between ice models).  0 (default): lanes continue their streams.  No
 * counterpart in the reference. */
int ppc_hip_seed_photons(ppc_hip_ctx* ctx, uint64_t seed);
/* read/write the per-lane MWC states (x | c << 32), lanes entries */
int ppc_hip_get_states(ppc_hip_ctx* ctx, uint64_t* states, int32_t lanes);
int ppc_hip_set_states(ppc_hip_ctx* ctx, const uint64_t* states, int32_t lanes);

/* start propagating num photons of nb bunches (kernel_c + kernel_f,
 * gpu/ppc.cu:285-306); returns without waiting for the kernel */
int ppc_hip_launch(ppc_hip_ctx* ctx, const ppc_hip_bunch* bunches, uint32_t nb, uint32_t num);
/* wait for the launch and read its counters (kernel_i, gpu/ppc.cu:255-277); PPC_HIP_E_INCOMPLETE
 * when a wave of the launch did not reach the kernel's end or a photon was not started
 * (the reference's check that every multiprocessor ran, gpu/ppc.cu:266) */
int ppc_hip_wait(ppc_hip_ctx* ctx, ppc_hip_stats* stats);
/* copy min(stored, cap) hits of the last launch to the host (gpu/ppc.cu:278-282);
 * photon (optional) receives the photon index of each hit */
int ppc_hip_hits(ppc_hip_ctx* ctx, ppc_hip_hit* hits, uint32_t* photon, uint32_t cap, uint32_t* n);
/* the replay records of the last launch */
int ppc_hip_records(ppc_hip_ctx* ctx, ppc_hip_record* rec, uint32_t cap, uint32_t* n);

/* ---- batches over ice variants ---- */
/* replace the per-layer factors by nvar ice variants, lay[v*size + layer]
 * (variant 0 first); variants share everything but the layer factors */
int ppc_hip_set_layers(ppc_hip_ctx* ctx, const ppc_hip_layer* lay, int32_t nvar);
/* like ppc_hip_launch, with bunch b propagated in ice variant variant[b] */
int ppc_hip_launch_batch(ppc_hip_ctx* ctx, const ppc_hip_bunch* bunches, const uint16_t* variant, uint32_t nb,
                         uint32_t num);
/* capture the launch of these bunches (counter reset and kernel; the bunches are
 * copied to buffers of the graph once) into a HIP graph; ppc_hip_graph_launch
 * replays it and ppc_hip_wait / ppc_hip_hits read its results as after
 * ppc_hip_launch.  The MWC streams continue, so every replay simulates the same
 * sources with new random numbers.  ppc_hip_set_layers drops the graph. */
int ppc_hip_graph_capture(ppc_hip_ctx* ctx, const ppc_hip_bunch* bunches, const uint16_t* variant, uint32_t nb,
                          uint32_t num);
int ppc_hip_graph_launch(ppc_hip_ctx* ctx);

/* ---- gradient of a Poisson log-likelihood of the hit counts ---- */
/* the photons of the last launch that hit (record_starts), in no particular order */
int ppc_hip_hit_photons(ppc_hip_ctx* ctx, ppc_hip_adj_record* rec, uint32_t cap, uint32_t* n);
/* Replay photons of a forward launch of these bunches (same bunches and num, ice
 * variant 0) and accumulate into grad[0..size) d(LLH)/d(absorption parameter
 * a_i) and into grad[size..2*size) d(LLH)/d(scattering parameter b_i).  The
 * photons are rec[0..nrec), or with rec == NULL the hitting photons recorded by
 * the last launch (record_starts; it must have been waited for).
 * weight[s*gsize + dom] is d(LLH)/d(mean hits of dom) for track segment s < nseg
 * (hits of later segments weigh 0); dpda[i] = d(p_i)/d(a_i) and
 * ds1db[i] = d(s1_i)/d(b_i) relate the parameters to the factorised layer table.
 * Each photon runs twice, summing its hits' weights and then scoring its path;
 * stats->photons counts both runs and stats->nonfinite the photons whose two runs
 * disagreed.  The results of the forward launch stay readable.  Synchronous. */
int ppc_hip_adjoint(ppc_hip_ctx* ctx, const ppc_hip_bunch* bunches, uint32_t nb, uint32_t num,
                    const ppc_hip_adj_record* rec, uint32_t nrec, const float* weight, uint32_t nseg,
                    const float* dpda, const float* ds1db, double* grad, ppc_hip_stats* stats);

/* ppc_hip_adjoint without device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight fitting
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as ppc_hip_adjoint;
 * its staging (records, weights, d(p)/d(a), d(s1)/d(b)), gradient accumulators and counters
 * live in a fixed scratch area of the GPU (2 MiB, set up by ppc_hip_open) instead of buffers
 * the context allocates on first use.  Limits: nseg * gsize <= 262144 weights and, when
 * rec is given, nrec <= 65536 records; larger calls fail with PPC_HIP_E_CAPACITY.  Every
 * call on one GPU uses the same scratch area (each call clears what must start at zero):
 * calls that overlap on one GPU, from two contexts or two threads, race. */
int ppc_hip_adjoint_noalloc(ppc_hip_ctx* ctx, const ppc_hip_bunch* bunches, uint32_t nb, uint32_t num,
                            const ppc_hip_adj_record* rec, uint32_t nrec, const float* weight, uint32_t nseg,
                            const float* dpda, const float* ds1db, double* grad, ppc_hip_stats* stats);

/* ---- hit post-processing on the device ---- */
/* sensor type: angular sensitivity and PMT layout (gpu/ini.cxx:122-281) */
typedef struct {
  int32_t def;        /* 1: one PMT with an angular-sensitivity polynomial; 0: PMTs along pmt_dirs */
  int32_t ns;         /* polynomial coefficients in s (def) */
  int32_t npmt, pmt0; /* PMTs of the type: pmt_dirs[pmt0 .. pmt0+npmt) (not def) */
  float mas;          /* def: maximum of the polynomial (<= 0: cut pos.tilt > s[0]) */
  float beta, fx, Rr, Rz, cable; /* not def: PMT acceptance curve, its normalisation, size, cable azimuth;
                                   beta -2: a wavelength-shifting tube of radius Rr and length
                                   -2 Rz (Rz < 0), read out at its two ends (gpu/ini.cxx:211-243) */
  float s[16];
} ppc_hip_sensor;

/* per DOM: what the acceptance of its hits needs (gpu/ini.cxx:86-98, gpu/f2k.cxx:266-324) */
typedef struct {
  int32_t sensor;   /* index into sensors, -1: the DOM records nothing (no high voltage or no acceptance) */
  int32_t acc;      /* row of rat */
  float rde;        /* the DOM's relative efficiency over the maximum of its acceptance class */
  float azi;        /* cable azimuth [deg], < 0: no cable */
  float tilt[3];    /* DOM axis */
  float R, F;       /* radius, axis ratio */
  float pad;
} ppc_hip_dom_accept;

typedef struct {
  int32_t nsensor;
  const ppc_hip_sensor* sensors;
  int32_t npmt;
  const float* pmt_dirs;          /* [npmt][3] */
  int32_t nacc;
  const float* rat;               /* [nacc][PPC_HIP_WNUM] acceptance of a wavelength bin */
  int32_t gsize;
  const ppc_hip_dom_accept* doms; /* [gsize] */
  const float* wvs;               /* [PPC_HIP_WNUM][2] low and high wavelength of a bin [nm] */
} ppc_hip_accept_tables;

/* the fate of one raw hit: accepted by PMT pmt (>= 0) or rejected (-1); for an accepted hit
 * the photon's wavelength [nm], the impact point relative to the DOM centre [m], the
 * photon direction (the reference's pout, gpu/f2k.cxx:340-352) and the time the sensor adds
 * to the hit [ns] (the transit through a wavelength-shifting tube, pmt 0 or 1 naming its end,
 * gpu/ini.cxx:211-243; 0 for other sensors) */
typedef struct {
  int32_t pmt;
  float wv;
  float r[3], n[3];
  float dt;
} ppc_hip_accepted;

/* copy the acceptance tables to the device (no counterpart: the reference applies them on
 * the host in print(), gpu/f2k.cxx:247-363) */
int ppc_hip_upload_accept(ppc_hip_ctx* ctx, const ppc_hip_accept_tables* tables);
/* decide acceptance, PMT and wavelength of the min(stored, cap) hits of the last launch on the
 * device, one record per hit in the order of ppc_hip_hits, with counter-based random numbers
 * of seed (statistically the reference's host draws, not the same numbers); synchronous */
int ppc_hip_accept(ppc_hip_ctx* ctx, uint64_t seed, ppc_hip_accepted* out, uint32_t cap, uint32_t* n);

/* message of the last failing call (thread local) */
const char* ppc_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
