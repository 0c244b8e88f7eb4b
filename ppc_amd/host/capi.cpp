// capi.cpp -- C facade of the xppc front-end (include/xppc_c.h).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "frontend.h"
#include "xppc_c.h"

using namespace xppc;

namespace {
bool g_init = false;
std::vector<int32_t> g_str, g_dom;
std::vector<float> g_wvs;
float g_eff;
int32_t g_ovr, g_l0;
}  // namespace

extern "C" {

int xppc_c_initialize(float enh) {
  initialize(enh);
  g_init = true;
  return 0;
}

const ppc_hip_tables* xppc_c_tables(void) {
  if (!g_init) return nullptr;
  det.refresh_pointers();
  return &det.t;
}

#define SCALAR(nm)                   \
  if (n == #nm) {                    \
    *bytes = sizeof(t.nm);           \
    return (const void*)&t.nm;       \
  }
#define ARRAY(nm, vec)                                   \
  if (n == nm) {                                         \
    *bytes = sizeof((vec)[0]) * (vec).size();            \
    return (const void*)(vec).data();                    \
  }

const void* xppc_c_field(const char* name, uint64_t* bytes) {
  if (!g_init || !name || !bytes) return nullptr;
  det.refresh_pointers();
  const ppc_hip_tables& t = det.t;
  std::string n(name);
  SCALAR(size) SCALAR(dh) SCALAR(rdh) SCALAR(hmin) SCALAR(ocv) SCALAR(sf) SCALAR(g) SCALAR(gr) SCALAR(xR)
  SCALAR(SF) SCALAR(G) SCALAR(GR) SCALAR(hr) SCALAR(hr2) SCALAR(hs) SCALAR(ha) SCALAR(hifl)
  SCALAR(azx) SCALAR(azy) SCALAR(k1) SCALAR(k2) SCALAR(kz) SCALAR(fr) SCALAR(sum) SCALAR(bfr)
  SCALAR(rx) SCALAR(cn) SCALAR(cl) SCALAR(crst) SCALAR(cb)
  SCALAR(tmod) SCALAR(vthk) SCALAR(lnum) SCALAR(lpts) SCALAR(lmin) SCALAR(lrdz) SCALAR(r0) SCALAR(lnx) SCALAR(lny)
  SCALAR(lr) SCALAR(mmin) SCALAR(mstp) SCALAR(mnum) SCALAR(mcut) SCALAR(mcol) SCALAR(nsc) SCALAR(gsize) SCALAR(rsize)
  ARRAY("lay", det.lay) ARRAY("az", det.az) ARRAY("wv", det.wv) ARRAY("is", det.is) ARRAY("ls", det.ls)
  ARRAY("sc", det.sc) ARRAY("lp", det.lp) ARRAY("oms", det.oms) ARRAY("rm", det.rm)
  ARRAY("abs_", det.abs_tab) ARRAY("sca", det.sca_tab)
  ARRAY("par_a", det.par_a) ARRAY("par_b", det.par_b) ARRAY("dpda", det.dpda) ARRAY("ds1db", det.ds1db)
  if (n == "str" || n == "dom") {
    g_str.resize(det.names.size());
    g_dom.resize(det.names.size());
    for (size_t i = 0; i < det.names.size(); i++) g_str[i] = det.names[i].str, g_dom[i] = det.names[i].dom;
    const std::vector<int32_t>& v = n == "str" ? g_str : g_dom;
    *bytes = sizeof(int32_t) * v.size();
    return v.data();
  }
  if (n == "wvs") {
    g_wvs.resize(3 * WNUM);
    for (int i = 0; i < WNUM; i++) g_wvs[3 * i] = det.wvs[i].w, g_wvs[3 * i + 1] = det.wvs[i].i, g_wvs[3 * i + 2] = det.wvs[i].f;
    *bytes = sizeof(float) * g_wvs.size();
    return g_wvs.data();
  }
  if (n == "eff") {
    g_eff = det.eff;
    *bytes = sizeof g_eff;
    return &g_eff;
  }
  if (n == "ovr") {
    g_ovr = (int32_t)det.ovr;
    *bytes = sizeof g_ovr;
    return &g_ovr;
  }
  if (n == "l0") {
    g_l0 = det.l0;
    *bytes = sizeof g_l0;
    return &g_l0;
  }
  *bytes = 0;
  return nullptr;
}

int xppc_c_extend_multipliers(int32_t lanes) {
  if (!g_init || det.rm.empty()) return -1;
  if ((int)det.rm.size() < lanes) {
    uint32_t lo = *std::min_element(det.rm.begin(), det.rm.end());
    size_t have = det.rm.size();
    det.rm.resize(lanes);
    int got = ppc_mwc_multipliers(lo, lanes - have, det.rm.data() + have);
    if (got != (int)(lanes - have)) {
      det.rm.resize(have + std::max(got, 0));
      return -1;
    }
  }
  det.refresh_pointers();
  return (int)det.rm.size();
}

int xppc_c_flasher_bunch(int32_t str, int32_t dom, ppc_hip_bunch* out) {
  if (!g_init || !out) return -1;
  *out = flasher_bunch(str, dom);
  out->num = 0;
  return 0;
}

double xppc_c_particle_bunch(float x, float y, float z, float t, float E, float dr, int32_t type, float nx, float ny,
                             float nz, ppc_hip_bunch* out) {
  if (!g_init || !out) return -1;
  double mean = particle_bunch(x, y, z, t, E, dr, type, nx, ny, nz, *out);
  out->num = 0;
  return mean;
}

int xppc_c_clst_bunch(float x, float y, float z, float t, float nx, float ny, float nz, float dr, float beta,
                      ppc_hip_bunch* out) {
  if (!g_init || !out) return -1;
  *out = clst_bunch(x, y, z, t, nx, ny, nz, dr, beta);
  out->num = 0;
  return 0;
}

double xppc_c_mopo_bunch(float x, float y, float z, float t, float nx, float ny, float nz, float light, float length,
                         float beta, float tau, float fr, ppc_hip_bunch* out) {
  if (!g_init || !out) return -1;
  double mean = mopo_bunch(x, y, z, t, nx, ny, nz, light, length, beta, tau, fr, *out);
  out->num = 0;
  return mean;
}

uint64_t xppc_c_flasher_photons(uint64_t num) {
  return (unsigned long long)(num * (det.eff / dppm) / det.ovr);
}

float xppc_c_zshift(float x, float y, float z, float* dz) {
  float r[4] = {x, y, z, 0};
  float h = det.t.dh;
  float s = host_zshift(det.t, r, h);
  if (dz) *dz = h;
  return s;
}

uint32_t xppc_c_seed(void) { return sv; }
float xppc_c_xrnd(void) { return xrnd(); }
float xppc_c_grnd(void) { return grnd(); }

int xppc_c_start(int32_t device) {
  if (!g_init) return -1;
  choose(device);
  ini();
  return 0;
}

int xppc_c_stop(void) {
  fin();
  stop();
  return 0;
}

int xppc_c_flset(int32_t str, int32_t dom) {
  flset(str, dom);
  return 0;
}

int xppc_c_flone(uint64_t num) {
  flone(num);
  return 0;
}

int xppc_c_sett(float nx, float ny, float nz, int32_t id, uint64_t sub, int32_t frame, double t0) {
  sett(nx, ny, nz, std::make_pair((int)id, (unsigned long long)sub), frame, t0);
  return 0;
}

int xppc_c_addp(float x, float y, float z, float t, float E, float dr, int32_t type, float scale) {
  addp(x, y, z, t, E, dr, type, scale);
  return 0;
}

int xppc_c_eout(void) {
  eout();
  return 0;
}

uint64_t xppc_c_hitz(int32_t* str, int32_t* dom, int32_t* pmt, float* t, float* wv, uint64_t cap) {
  uint64_t n = 0;
  for (const auto& e : hitz)
    for (const pout& po : e.second) {
      if (str && n < cap) {
        str[n] = e.first.omkey.str, dom[n] = e.first.omkey.dom, pmt[n] = e.first.pmt;
        t[n] = po.r[3], wv[n] = po.n[3];
      }
      n++;
    }
  return n;
}

uint64_t xppc_c_hitz_keys(float* time, float* dir, uint32_t* count, uint64_t cap) {
  uint64_t n = 0;
  for (const auto& e : hitz) {
    if (time && n < cap) time[n] = e.first.time, dir[n] = e.first.dir, count[n] = (uint32_t)e.second.size();
    n++;
  }
  return n;
}

uint64_t xppc_c_hitz_tracks(int32_t* frame, int32_t* id, uint64_t* sub, double* t0, uint64_t cap) {
  uint64_t n = 0;
  for (const auto& e : hitz) {
    if (frame && n < cap) {
      const mcid& m = e.first.track;
      frame[n] = m.frame, id[n] = m.first, sub[n] = m.second, t0[n] = m.t0;
    }
    n++;
  }
  return n;
}

int xppc_c_flshift(const float* m, float* n, float* r) {
  float mm[3] = {m[0], m[1], m[2]};
  flshift(mm, n, r);
  return 0;
}

int xppc_c_set_res(float t, float c) {
  set_res(t, c);
  return 0;
}

int xppc_c_efin(void) {
  efin();
  return 0;
}

const ppc_hip_accept_tables* xppc_c_accept_tables(void) { return g_init ? accept_tables() : nullptr; }

int xppc_c_totals(uint64_t* photons, uint64_t* hits, double* device_ms) {
  BackendInfo b = backend_info();
  if (photons) *photons = b.photons;
  if (hits) *hits = b.hits;
  if (device_ms) *device_ms = b.device_ms;
  return 0;
}

}  // extern "C"
