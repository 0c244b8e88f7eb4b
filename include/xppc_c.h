/*
 * xppc_c.h -- C facade of libxppc.so for foreign-function callers.
 *
 * The reference exposes its front-end only as the C++ namespace xppc
 * (llh/ppc.h); this facade gives the same operations plain C signatures so that
 * ctypes/cgo/JNI-style bindings can build the tables from an ice directory,
 * make the photon bunches the front-ends make, and hand both to the device
 * library (ppc_hip.h).  Every function refers to the single process-wide
 * detector that xppc::initialize() builds.
 */
#ifndef XPPC_C_H
#define XPPC_C_H

#include <stdint.h>

#include "ppc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* the first n multipliers a < `below` with a*2^32-1 and (a*2^32-2)/2 prime, in
 * descending order (the definition of the ice directories' rnd.txt); returns n */
int ppc_mwc_multipliers(uint32_t below, uint32_t n, uint32_t* out);

/* xppc::initialize (gpu/ppc.cu:36): read the ice directory named by the
 * PPCTABLESDIR / ICEMODELDIR / ... environment (gpu/ini.cxx:584-625) */
int xppc_c_initialize(float enh);

/* the device tables of the initialised detector */
const ppc_hip_tables* xppc_c_tables(void);

/* pointer to and size of a named table of the initialised detector:
 * scalars and arrays of ppc_hip_tables by member name, plus
 *   abs_, sca     dense [512][172] ice tables as gpu/ini.cxx:1684-1691 builds them
 *   str, dom      string and DOM number of every in-ice DOM (int32)
 *   wvs           [512][3] wavelength bin centre, low and high edge [nm]
 *   eff, ovr, l0  light-yield scale, oversize factor xR^2, zero string of the 1-d tilt
 *   par_a, par_b  per layer (layer 0 deepest) the icemodel.dat parameters adust(400), be(400)
 *   dpda, ds1db   per layer d(lay.p)/d(par_a) and d(lay.s1)/d(par_b) (float) */
const void* xppc_c_field(const char* name, uint64_t* bytes);

/* lengthen the multiplier table to at least `lanes` entries, continuing the
 * safe-prime sequence below the smallest multiplier present; returns rsize */
int xppc_c_extend_multipliers(int32_t lanes);

/* the flasher bunch flset(str, dom) configures (gpu/f2k.cxx:630-759);
 * num is left at 0 */
int xppc_c_flasher_bunch(int32_t str, int32_t dom, ppc_hip_bunch* out);

/* the bunch addp() builds for a particle of f2k type (0 muon, 11-14 EM,
 * 101-107 hadron; gpu/f2k.cxx:17-167,484-504) with direction n; returns the
 * Poisson mean of its photon count, num is left at 0 */
double xppc_c_particle_bunch(float x, float y, float z, float t, float E, float dr, int32_t type, float nx,
                             float ny, float nz, ppc_hip_bunch* out);

/* the bunches of the IceTray entry points with direction n (gpu/f2k.cxx:507-528): addp_clst's
 * segment of length dr of a particle of velocity beta, and addp_mopo's segment with light
 * yield `light`, luminescence time tau and bare-track fraction fr (returns the Poisson mean of
 * its photon count); num is left at 0 */
int xppc_c_clst_bunch(float x, float y, float z, float t, float nx, float ny, float nz, float dr, float beta,
                      ppc_hip_bunch* out);
double xppc_c_mopo_bunch(float x, float y, float z, float t, float nx, float ny, float nz, float light, float length,
                         float beta, float tau, float fr, ppc_hip_bunch* out);

/* photons flone(num) propagates for a flasher pulse: num*eff/dppm/xR^2 (gpu/f2k.cxx:586-588) */
uint64_t xppc_c_flasher_photons(uint64_t num);

/* depth shift of the ice layers at (x,y,z) and the local layer thickness (gpu/pro.cu:168-264) */
float xppc_c_zshift(float x, float y, float z, float* dz);

/* the host random-number state (seed of the MWC streams, gpu/ini.cxx:554) */
uint32_t xppc_c_seed(void);
/* the host generator's next uniform in (0, 1] and gaussian (xppc::xrnd, xppc::grnd; gpu/f2k.cxx:1-10) */
float xppc_c_xrnd(void);
float xppc_c_grnd(void);

/* the library-mode pipeline of the xppc interface (what llh drives, llh/ppc.h):
 * choose + ini, flasher set-up, track segments, particles, flush, and the hitz
 * map flattened into arrays (one entry per recorded photon, map order) */
int xppc_c_start(int32_t device);     /* choose(device); ini() */
int xppc_c_stop(void);                /* fin(); stop() */
int xppc_c_flset(int32_t str, int32_t dom);
int xppc_c_flone(uint64_t num);
int xppc_c_sett(float nx, float ny, float nz, int32_t id, uint64_t sub, int32_t frame, double t0);
int xppc_c_addp(float x, float y, float z, float t, float E, float dr, int32_t type, float scale);
int xppc_c_eout(void);
/* number of photon records in hitz; with non-NULL arrays, fills up to cap of them */
uint64_t xppc_c_hitz(int32_t* str, int32_t* dom, int32_t* pmt, float* t, float* wv, uint64_t cap);
/* number of keys of hitz; with non-NULL arrays, fills up to cap of them in key order: the key's
 * (binned) time and direction and the number of photon records under it */
uint64_t xppc_c_hitz_keys(float* time, float* dir, uint32_t* count, uint64_t cap);
/* number of keys of hitz; with non-NULL arrays, fills up to cap of them in key order: the
 * segment each key belongs to (the mcid sett() gave it: frame, id, sub-id, t0) */
uint64_t xppc_c_hitz_tracks(int32_t* frame, int32_t* id, uint64_t* sub, double* t0, uint64_t cap);
/* flshift(m, n, r): turn n (and r unless NULL) by the rotation that takes the z axis onto the
 * DOM axis m; n with a NaN z first gets the LED elevation of the last flset (gpu/f2k.cxx:590-628) */
int xppc_c_flshift(const float* m, float* n, float* r);
/* set_res(t, c): time and direction bins of the hitz keys (gpu/f2k.cxx:200-214) */
int xppc_c_set_res(float t, float c);
int xppc_c_efin(void);
/* the acceptance tables of print() in the layout the device post-processing takes
 * (ppc_hip_upload_accept); valid until the next call */
const ppc_hip_accept_tables* xppc_c_accept_tables(void);
/* totals of the launches since start: photons propagated, raw hits, device time [ms] */
int xppc_c_totals(uint64_t* photons, uint64_t* hits, double* device_ms);

#ifdef __cplusplus
}
#endif

#endif
