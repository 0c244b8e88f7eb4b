// kparams.h -- the parameter block of the propagation kernel.
//
// The block lives in device memory and the kernel reads it through a
// `const KParams*` argument (uniform fields become scalar loads, repeated where
// the compiler needs them again; passed by value, its ~600 B would occupy SGPRs
// the kernel needs).  The small per-layer, per-string and cell tables are packed
// by the host into one contiguous image that every workgroup copies into LDS at
// start; the large tables (DOM positions, the 1-d tilt table, a 2-d tilt table that
// does not fit in LDS, wavelength factors, the clearance map) stay in HBM and are
// reached through L2.
#ifndef PPC_KPARAMS_H
#define PPC_KPARAMS_H

#include <stdint.h>

#include "ppc_hip.h"

namespace ppc {

constexpr int kLMAX = PPC_HIP_LMAX;
constexpr int kWNUM = PPC_HIP_WNUM;
constexpr int kCX = PPC_HIP_CX, kCY = PPC_HIP_CY;
constexpr int kCTY = PPC_HIP_CTY;
// Waves per SIMD the propagation kernel is compiled for (the register budget: 72 VGPRs at
// 7 waves); the default launch fills every SIMD with that many waves
constexpr int kWavesPerSimd = 7;

// global counters of one launch
struct Counters {
  unsigned long long photons, steps, hits, killed, nonfinite;
  unsigned long long crossings, strings, doms, scatters;
  unsigned int next;   // next photon to hand out (dynamic refill)
  unsigned int hidx;   // next free hit slot
  unsigned int nphot;  // photons with hits recorded for the adjoint
  unsigned int waves;  // waves that reached the end of the kernel: all of the launch's
  unsigned int pad;
  // wall clock (wall_clock64 ticks): the launch's earliest wave start and its earliest and
  // latest wave end; the first two complemented, so that atomicMax on the zeroed block finds them
  unsigned long long t_start_c, t_end_min_c, t_end_max;
};

struct KParams {
  // ---- ice / detector scalars (names follow gpu/ini.cxx:452-518) ----
  int size;
  float dh, rdh, hmin;
  float ocv, sf, g, gr, xR;
  float SF, G, GR, hr, hr2, hs, ha, hifl;
  float azx, azy, k1, k2, kz, fr;
  float sum, bfr[12];
  float rx;
  int cn0, cn1;
  float cl0, cl1, crst0, crst1;
  float cb00, cb01, cb10, cb11;
  int tmod, vthk, lnum, lpts;
  float lmin, lrdz, r0, lnx, lny;
  float lr[kLMAX];  // 1-d tilt reference distances (also in the LDS image for per-lane indexing)
  float mmin0, mmin1, mstp0, mstp1;
  int mnum0, mnum1, mcut0, mcut1;
  int nsc, gsize;
  int uniform_dom;  // every DOM has R = OMR and F = 1
  int scat_frame;   // some layer has k1 != 1 or k2 != 1: scatter in the anisotropic frame
  // scattering-function constants: 1/(1-sf), 1-g^2, (1-g)^2, 1/(2g), 1/sf; [0] bulk ice, [1] hole ice
  float hg_isf1[2], hg_a[2], hg_b[2], hg_i2g[2], hg_isf[2];

  // ---- LDS image (byte offsets into it) ----
  const uint4* image;
  int image_words;  // 16-byte words
  int off_lay, off_az, off_lr, off_mcol, off_is, off_ls, off_sc, off_lp;
  int lp_in_lds, lp_stride;  // 1-D tilt table in LDS; row stride in float2

  // ---- HBM tables ----
  const float2* lp;         // [lnum][LYRS] (value, slope), used when not in LDS
  const ppc_hip_wv* wv;     // [512]
  const float4* dom;        // [gsize] x, y, z, R
  const float* domF;        // [gsize]
  // clearance map: per square of side 1/cm_rstep over the strings' footprint, a lower
  // bound of the horizontal distance from any point of the square to the nearest
  // string cylinder (radius of its line, or of the hole ice if larger)
  const float* clear_map;   // [cm_ny][cm_nx], or nullptr
  float cm_x0, cm_y0, cm_rstep;
  int cm_nx, cm_ny;

  // ---- RNG ----
  const uint32_t* rm;  // multiplier per lane
  uint2* rs;           // (x, c) per lane

  // ---- launch ----
  const ppc_hip_bunch* pz;
  const uint32_t* pstart;  // exclusive prefix sums of bunch sizes, [nb+1]
  uint32_t nb, num;
  ppc_hip_hit* hits;
  uint32_t* hit_photon;
  uint32_t hnum;
  Counters* ctr;
  ppc_hip_record* rec;
  uint32_t rec_stride, rec_max;
  int refill, deterministic;
  uint32_t chunk;  // photons claimed per atomic on the global counter
  uint64_t photon_seed;  // != 0: each photon starts from a state hashed from (photon_seed, index)
  const uint16_t* bvar;  // ice variant of each bunch (nullptr: variant 0)
  ppc_hip_adj_record* hit_photons;  // start state of every photon that hits, or nullptr
  uint32_t hit_photons_cap;
  // adjoint replay
  const ppc_hip_adj_record* adj_rec;
  uint32_t adj_nrec, adj_nseg;
  const float* adj_weight;  // [adj_nseg][gsize] d(LLH)/d(hits of a DOM)
  const float* adj_dpda;    // [size] d(p)/d(absorption parameter)
  const float* adj_ds1db;   // [size] d(s1)/d(scattering parameter)
  double* adj_grad;         // [2][size]
  uint32_t lanes;
};

// tables of the hit post-processing kernel (accept.hip)
struct AcceptParams {
  const ppc_hip_sensor* sensors;
  const float* pmt_dirs;
  const float* rat;
  const ppc_hip_dom_accept* doms;
  const float* wvs;
  int32_t gsize;
};

}  // namespace ppc

#endif
