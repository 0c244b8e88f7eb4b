// frontend.cpp -- the xppc front-end: light yield, photon bunching, the
// device backend seam and hit post-processing.
//
// Restates gpu/f2k.cxx (sources, bunching, print) and the backend seam of
// gpu/ppc.cu:38-509 on top of the C-ABI of libppc_hip.so: one GPU, one launch
// in flight while the host post-processes the previous one (gpu/ppc.cu:432-469).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <thread>
#include <tuple>
#include <iostream>
#include <sstream>
#include <string>

#include "frontend.h"
#include "xppc_c.h"

using std::cerr;
using std::endl;
using std::string;

namespace xppc {

// float arguments select the float overloads, as in the reference (gpu/ppc.cu:26)
using std::acos;
using std::atan2;
using std::cos;
using std::exp;
using std::floor;
using std::log;
using std::pow;
using std::sin;
using std::sqrt;
using std::tan;

bool cli_mode = false;
outz hitz;

// ------------------------------------------------------------ host RNG

float xrnd() {  // uniform in (0,1) from rand_r on the shared seed (gpu/f2k.cxx:1-6)
  unsigned int rnd;
  do rnd = rand_r(&sv);
  while (rnd == 0);
  const float a = 1.0f / (1ll + RAND_MAX);
  return a * rnd;
}

float grnd() {  // gpu/f2k.cxx:8-10
  float r = sqrtf(-2 * logf(xrnd()));
  return r * sinf(2 * FPI * xrnd());
}

// ------------------------------------------------------------ backend seam

namespace {

struct Backend {
  ppc_hip_ctx* ctx = nullptr;
  int device = -1;
  int lanes = 0, threads = 0;
  unsigned pmax = 0, pmxo = 0, hnum = 0, hquo = 1;
  std::vector<ppc_hip_bunch> pz;
  std::vector<ppc_hip_hit> hits;
  std::vector<ppc_hip_accepted> acc;  // device acceptance of hits (PPC_ACCEPT=device)
  bool dev_accept = false;
  unsigned long long launches = 0;
  unsigned hidx = 0;
  unsigned old = 0;
  double device_ms = 0;
  double span_min_ms = 0, span_max_ms = 0;  // summed in-kernel spans (gpu/ppc.cu:267-273)
  unsigned long long photons = 0, nhits = 0, steps = 0;
};

Backend be;
unsigned int pn = 0, pk = 0;
ppc_hip_bunch p{}, pfl{};

void check(int rc, const char* what) {
  if (rc == PPC_HIP_E_INCOMPLETE) {  // gpu/ppc.cu:266
    cerr << "Error: " << ppc_hip_last_error() << endl;
    exit(4);
  }
  if (rc != PPC_HIP_OK) {
    cerr << "HIP Error (" << what << "): " << ppc_hip_last_error() << endl;
    exit(2);
  }
}

// segment bookkeeping (gpu/f2k.cxx:216-231)
std::deque<mcid> flnz_lib;
std::deque<string> flnz_cli;
unsigned int flnb = 0, flne = 0, flnd = 0;
ihit tmph;
float trz = 0, crz = 0;

void pop_segment() {  // segments without an entry (no sett() call) keep the previous track id
  if (cli_mode) {
    if (flnz_cli.empty()) return;
    std::cout << flnz_cli.front() << "\n";
    flnz_cli.pop_front();
  } else {
    if (flnz_lib.empty()) return;
    tmph.track = flnz_lib.front();
    flnz_lib.pop_front();
  }
}

float res(float val, float bin) { return bin > 0 ? (floor(val / bin) + 0.5) * bin : val; }

bool deterministic() {
  const char* e = getenv("DTMN");
  return e != nullptr && atoi(e) != 0;
}

}  // namespace

// the acceptance of print() on the device (PPC_ACCEPT=device): sensor types, acceptance
// classes and per-DOM data in the layout of ppc_hip_accept_tables (valid until the next call)
const ppc_hip_accept_tables* accept_tables() {
  static std::vector<ppc_hip_sensor> sensors;
  static std::vector<float> dirs, rat, wvs;
  static std::vector<ppc_hip_dom_accept> doms;
  static ppc_hip_accept_tables at;
  sensors.clear(), dirs.clear(), rat.clear(), wvs.clear(), doms.clear();
  std::map<int, int> sidx;
  for (const auto& kv : det.types) {
    const SensorType& S = kv.second;
    ppc_hip_sensor x{};
    x.def = S.def && S.beta != -2 ? 1 : 0;
    x.beta = S.beta, x.fx = S.fx, x.Rr = S.Rr, x.Rz = S.Rz, x.cable = S.cable, x.mas = S.mas;
    if (x.def) {
      if (S.s.size() > 16) {
        cerr << "Error: device acceptance takes at most 16 angular sensitivity coefficients" << endl;
        exit(1);
      }
      x.ns = (int)S.s.size();
      for (size_t k = 0; k < S.s.size(); k++) x.s[k] = S.s[k];
    } else {
      x.pmt0 = (int)(dirs.size() / 3), x.npmt = (int)S.dirs.size();
      for (const Vec3& d : S.dirs) dirs.push_back(d[0]), dirs.push_back(d[1]), dirs.push_back(d[2]);
    }
    sidx[kv.first] = (int)sensors.size();
    sensors.push_back(x);
  }
  std::map<std::pair<int, int>, int> aidx;
  for (const auto& kv : det.accept)
    if (kv.second.rde > 0 && (int)kv.second.rat.size() == WNUM) {
      aidx[kv.first] = (int)(rat.size() / WNUM);
      rat.insert(rat.end(), kv.second.rat.begin(), kv.second.rat.end());
    }
  doms.assign(det.t.gsize, ppc_hip_dom_accept{});
  for (int i = 0; i < det.t.gsize; i++) {
    const DomInfo& n = det.names[i];
    ppc_hip_dom_accept& d = doms[i];
    d.sensor = -1, d.acc = 0;
    auto a = aidx.find(std::make_pair(n.omt, n.type));
    auto t = sidx.find(n.omt);
    if (n.hv > 0 && a != aidx.end() && t != sidx.end()) {
      d.sensor = t->second, d.acc = a->second;
      d.rde = n.rde / det.accept[a->first].rmax;
    }
    d.azi = n.azi;
    for (int k = 0; k < 3; k++) d.tilt[k] = n.tilt[k];
    d.R = det.oms[i].R, d.F = det.oms[i].F;
  }
  wvs.assign(2 * WNUM, 0.f);
  for (int k = 0; k < WNUM; k++) wvs[2 * k] = det.wvs[k].i, wvs[2 * k + 1] = det.wvs[k].f;
  at = ppc_hip_accept_tables{};
  at.nsensor = (int)sensors.size(), at.sensors = sensors.data();
  at.npmt = (int)(dirs.size() / 3), at.pmt_dirs = dirs.data();
  at.nacc = (int)(rat.size() / WNUM), at.rat = rat.data();
  at.gsize = det.t.gsize, at.doms = doms.data();
  at.wvs = wvs.data();
  return &at;
}

void print();

void start() {}

void stop() {
  fprintf(stderr, "\nDevice time: %2.1f (in-kernel: %2.1f...%2.1f) [ms]\n", be.device_ms, be.span_min_ms,
          be.span_max_ms);
  if (be.ctx) ppc_hip_close(be.ctx), be.ctx = nullptr;
}

void list_devices() {
  int32_t n = 0;
  if (ppc_hip_device_count(&n) != PPC_HIP_OK) n = 0;
  fprintf(stderr, "Found %d devices\n", n);
  for (int i = 0; i < n; i++) {
    char name[256];
    int32_t cus = 0;
    uint64_t mem = 0;
    if (ppc_hip_device_info(i, name, sizeof name, &cus, &mem) == PPC_HIP_OK)
      fprintf(stderr, "%d: %s CUs(%d) G(%llu)\n", i, name, cus, (unsigned long long)mem);
  }
  fprintf(stderr, "\n");
}

void choose(int device) {  // gpu/ppc.cu:490-508 (one GPU)
  if (device < 0) device = 0;
  if (be.ctx) return;
  sv += device;
  ppc_hip_config cfg{};
  if (const char* e = getenv("PPC_BLOCKS")) cfg.blocks = atoi(e);
  if (const char* e = getenv("PPC_THREADS")) cfg.threads = atoi(e);
  if (const char* e = getenv("PPC_REFILL")) cfg.refill = atoi(e);
  cfg.deterministic = deterministic() ? 1 : 0;
  int rc = ppc_hip_open(device, &cfg, &be.ctx);
  if (rc != PPC_HIP_OK) {
    cerr << "No active GPU(s) selected! (" << ppc_hip_last_error() << ")" << endl;
    exit(5);
  }
  be.device = device;
  check(ppc_hip_lanes(be.ctx, &be.lanes), "lanes");
  be.threads = cfg.threads > 0 ? cfg.threads : 256;
  cerr << "Running on device " << device << " with " << be.lanes << " lanes" << endl;
}

void ini() {  // gpu/ppc.cu:330-361 and gpu::ini 151-237
  if (!be.ctx) choose(0);
  {
    const char* e = getenv("HQUO");
    be.hquo = e == nullptr ? 1 : atoi(e);
    if (be.hquo < 1) be.hquo = 1;
    cerr << "HQUO(photons/max number of hits)=" << be.hquo << endl;
  }
  // photons per lane per launch (NPHO, gpu/ini.cxx:21; the reference's default is 1024): with
  // 458752 lanes, 512 makes launches of 2.3e8 photons, so that the host post-processing of
  // one launch (print(), ~0.3 us per hit) finishes within the next launch's kernel time
  unsigned npho = 512;
  {  // NPHO_<device>, else NPHO (gpu/ppc.cu:120-131); a value <= 0 leaves no device to run on
    const std::string key = "NPHO_" + std::to_string(be.device);
    const char* e = getenv(key.c_str());
    if (e == nullptr) e = getenv("NPHO");
    if (e != nullptr && *e != 0) {
      const int v = atoi(e);
      cerr << "Setting NPHO=" << v << endl;
      if (v <= 0) {
        cerr << "Not using device # " << be.device << "!" << endl;
        cerr << "No active GPU(s) selected!" << endl;
        exit(5);
      }
      npho = (unsigned)v;
    }
  }
  unsigned long long pm = (unsigned long long)be.lanes * npho;
  if (pm > 0xffffff00ULL) pm = 0xffffff00ULL;
  be.pmax = (unsigned)pm;
  be.pmxo = be.pmax / 128;  // at least OVER=128 photons per bunch on the GPU
  if (be.pmxo < 1) be.pmxo = 1;
  be.pmax = be.pmxo * 128;
  uint32_t cap_b = 0, cap_h = 0;
  check(ppc_hip_capacity(be.ctx, &cap_b, &cap_h), "capacity");
  if (be.pmxo > cap_b) be.pmxo = cap_b;
  be.hnum = be.pmax / be.hquo;
  if (be.hnum > cap_h) be.hnum = cap_h;

  // one MWC stream per lane: continue the safe-prime sequence below the file's smallest multiplier
  if ((int)det.rm.size() < be.lanes) {
    uint32_t lo = *std::min_element(det.rm.begin(), det.rm.end());
    size_t have = det.rm.size();
    det.rm.resize(be.lanes);
    int got = ppc_mwc_multipliers(lo, be.lanes - have, det.rm.data() + have);
    if (got != (int)(be.lanes - have)) {
      cerr << "Error: not enough multipliers: only have " << have + got << " (need " << be.lanes << ")!" << endl;
      exit(1);
    }
    cerr << "Extended the multiplier table from " << have << " to " << be.lanes << endl;
  }
  det.refresh_pointers();
  check(ppc_hip_upload(be.ctx, &det.t), "upload");
  check(ppc_hip_seed(be.ctx, sv), "seed");
  be.pz.assign(be.pmxo, ppc_hip_bunch{});
  be.hits.assign(be.hnum, ppc_hip_hit{});
  {
    const char* e = getenv("PPC_ACCEPT");
    be.dev_accept = e != nullptr && std::string(e) == "device";
  }
  if (be.dev_accept) {
    check(ppc_hip_upload_accept(be.ctx, accept_tables()), "acceptance tables");
    be.acc.assign(be.hnum, ppc_hip_accepted{});
    cerr << "Hit acceptance on the device" << endl;
  }
  be.old = 0;
  pn = 0, pk = 0;
  cerr << "Photons per launch " << be.pmax << ", bunches " << be.pmxo << ", hits " << be.hnum << endl;
}

void fin() {
  be.pz.clear();
  be.hits.clear();
}

size_t getMaxBunchSize() { return be.pmxo; }
size_t getWorkgroupSize() { return be.threads; }
float getTotalDeviceTime() { return (float)be.device_ms; }

BackendInfo backend_info() {
  return BackendInfo{be.lanes, be.threads, be.pmax, be.pmxo, be.hnum, be.device_ms, be.photons, be.nhits, be.steps};
}

// launch num photons of the current bunches; post-process the previous launch (gpu/ppc.cu:432-469)
static void kernel(unsigned num) {
  if (be.old > 0) {
    ppc_hip_stats st{};
    check(ppc_hip_wait(be.ctx, &st), "wait");
    be.device_ms += st.device_ms;
    be.span_min_ms += st.span_min_ms, be.span_max_ms += st.span_max_ms;
    be.photons += st.photons, be.nhits += st.hits, be.steps += st.steps;
    if (st.nonfinite > 0) cerr << "Error: TOT was a nan or an inf " << st.nonfinite << " times!" << endl;
    if (st.hits > be.hnum) cerr << "Error: data buffer overflow occurred: " << st.hits << ">" << be.hnum << "!" << endl;
    uint32_t n = 0;
    check(ppc_hip_hits(be.ctx, be.hits.data(), nullptr, be.hnum, &n), "hits");
    be.hidx = n;
    if (be.dev_accept) {  // before the next launch reuses the hit buffer
      uint32_t m = 0;
      check(ppc_hip_accept(be.ctx, ((unsigned long long)sv << 32) ^ be.launches, be.acc.data(), be.hnum, &m), "accept");
      be.launches++;
    }
    cerr << "photons: " << be.old << "  hits: " << be.hidx << endl;
  }
  if (num > 0) check(ppc_hip_launch(be.ctx, be.pz.data(), pk, num), "launch");
  if (be.old > 0) print();
  be.old = num;
}

static void output() {
  kernel(pn);
  pn = 0, pk = 0;
  flnd = flne;
}

static void addh(unsigned long long num) {  // split into bunches and launches (gpu/f2k.cxx:377-385)
  do {
    p.num = (uint32_t)std::min((unsigned long long)(be.pmax - pn), num);
    num -= p.num;
    pn += p.num;
    be.pz[pk++] = p;
    if (pn == be.pmax || pk == be.pmxo) output();
  } while (num > 0);
  if (pn > 0 && pk == 0) be.pz[pk++] = p;
}

// ------------------------------------------------------------ light yield

static const float rho = 0.9216f;      // ice density [mwe]
static const float m0 = 0.105658389f;  // muon mass [GeV]

// photons per GeV / per meter and the longitudinal profile of each particle type (gpu/f2k.cxx:17-167)
static float yield(float E, float dr, int type) {
  p.n[3] = dr;
  float nph = 0;
  if (type < 0)
    nph = E / dppm;
  else {
    float logE = logf(std::max(m0, E));
    if (type < 10) {  // bare muon (0) or its end point (1)
      float extr = 1 + std::max(0.0f, 0.1880f + 0.0206f * (logE - type)) * 0.910f / rho;
      nph = dr > 0 ? dr * extr : 0;
      p.f = 1 / extr;
      p.u.p.a = 0, p.u.p.b = 0;
    } else {  // cascades
      const float Lrad = 0.39652 * 0.910f / rho;
      const float em = 5.321 * 0.910f / rho;
      nph = em * E;
      if (type < 100) {
        switch (type) {
          default:
          case 10:
          case 11:
          case 12: p.u.p.a = 2.01849f + 0.63176f * logE, p.u.p.b = Lrad / 0.63207f; break;
          case 13: p.u.p.a = 2.00035f + 0.63190f * logE, p.u.p.b = Lrad / 0.63008f; break;
          case 14: p.u.p.a = 2.83923f + 0.58209f * logE, p.u.p.b = Lrad / 0.64526f; break;
        }
      } else {
        float E0, m, f0, rms0, gamma;
        switch (type) {
          default:
          case 101:
          case 102:
            p.u.p.a = 1.58357292f + 0.41886807f * logE, p.u.p.b = Lrad / 0.33833116f;
            E0 = 0.18791678f, m = 0.16267529f, f0 = 0.30974123f, rms0 = 0.95899551f, gamma = 1.35589541f;
            break;
          case 103:
            p.u.p.a = 1.69176636f + 0.40803489f * logE, p.u.p.b = Lrad / 0.34108075f;
            E0 = 0.19826506f, m = 0.16218006f, f0 = 0.31859323f, rms0 = 0.94033488f, gamma = 1.35070162f;
            break;
          case 104:
            p.u.p.a = 1.95948974f + 0.34934666f * logE, p.u.p.b = Lrad / 0.34535151f;
            E0 = 0.21687243f, m = 0.16861530f, f0 = 0.27724987f, rms0 = 1.00318874f, gamma = 1.37528605f;
            break;
          case 105:
            p.u.p.a = 1.47495778f + 0.40450398f * logE, p.u.p.b = Lrad / 0.35226706f;
            E0 = 0.29579368f, m = 0.19373018f, f0 = 0.02455403f, rms0 = 1.01619344f, gamma = 1.45477346f;
            break;
          case 106:
            p.u.p.a = 1.57739060f + 0.40631102f * logE, p.u.p.b = Lrad / 0.35269455f;
            E0 = 0.66725124f, m = 0.19263595f, f0 = 0.17559033f, rms0 = 1.01414337f, gamma = 1.45086895f;
            break;
          case 107:
            p.u.p.a = 1.92249171f + 0.33701751f * logE, p.u.p.b = Lrad / 0.34969748f;
            E0 = 0.29579368f, m = 0.19373018f, f0 = 0.02455403f, rms0 = 1.01094637f, gamma = 1.50438415f;
            break;
        }
        float e = std::max(2.71828183f, E), f;
        float F = 1 - powf(e / E0, -m) * (1 - f0);
        float dF = F * rms0 * powf(logf(e), -gamma);
        do f = F + dF * grnd();
        while (f < 0 || 1.1 < f);
        nph *= f;
      }
      p.f = 0;
      if (dr > 0) p.u.p.a = 0, p.u.p.b = 0;
    }
  }
  return det.eff * nph;
}

static double gammln(double xx) {  // log Gamma, Lanczos series (gpu/f2k.cxx:387-401)
  static const double stp = 2.5066282746310005;
  static const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                -1.231739572450155, 0.1208650973866179e-2, -0.5395239384953e-5};
  double ser = 1, x = xx - 1, tmp = x + 5.5;
  tmp -= (x + 0.5) * log(tmp);
  for (int j = 0; j <= 5; j++) {
    x++;
    ser += cof[j] / x;
  }
  return -tmp + log(stp * ser);
}

static unsigned long long poidev(double xm) {  // Poisson: product method / rejection (gpu/f2k.cxx:403-433)
  double sq, alxm, g, t, y;
  unsigned long long um;
  if (xm < 12) {
    g = exp(-xm);
    um = -1ULL;
    t = 1;
    do {
      um++;
      t *= xrnd();
    } while (t > g);
  } else {
    sq = sqrt(2 * xm);
    alxm = log(xm);
    g = xm * alxm - gammln(xm + 1);
    do {
      double em;
      do {
        y = tan(FPI * xrnd());
        em = sq * y + xm;
      } while (em < 0);
      um = floor(em);
      t = 0.9 * (1 + y * y) * exp(um * alxm - gammln(um + 1) - g);
    } while (xrnd() > t);
  }
  return um;
}

static unsigned long long bnldev(unsigned long long n, double pp) {  // binomial (gpu/f2k.cxx:435-476)
  unsigned long long j, bnl;
  double am, em, g, angle, pr, sq, t, y, pc, plog, pclog, en, oldg;
  pr = (pp <= 0.5 ? pp : 1 - pp);
  am = n * pr;
  if (n < 25) {
    bnl = 0;
    for (j = 1; j <= n; j++)
      if (xrnd() < pr) bnl++;
  } else if (am < 1) {
    g = exp(-am);
    t = 1;
    for (j = 0; j <= n; j++) {
      t *= xrnd();
      if (t < g) break;
    }
    bnl = (j <= n ? j : n);
  } else {
    en = n;
    oldg = gammln(en + 1);
    pc = 1 - pr;
    plog = log(pr);
    pclog = log(pc);
    sq = sqrt(2 * am * pc);
    do {
      do {
        angle = FPI * xrnd();
        y = tan(angle);
        em = sq * y + am;
      } while (em < 0 || em >= (en + 1));
      em = floor(em);
      t = 1.2 * sq * (1 + y * y) * exp(oldg - gammln(em + 1) - gammln(en - em + 1) + em * plog + (en - em) * pclog);
    } while (xrnd() > t);
    bnl = em;
  }
  if (pr != pp) bnl = n - bnl;
  return bnl;
}

// ------------------------------------------------------------ sources

static void addf(float rx, float ry, float rz, float t, unsigned long long num) {  // gpu/f2k.cxx:478-482
  p.r[0] = rx, p.r[1] = ry, p.r[2] = rz, p.r[3] = t;
  p.type = 1;
  p.u.l.ka = 0, p.u.l.up = 0, p.u.l.fldr = -1, p.u.l.fla = -1, p.u.l.ofla = -1;
  p.n[3] = 1;
  addh(num);
}

// set p for a particle (or a pre-configured flasher for type < 0); returns the mean photon count
static double particle_setup(float rx, float ry, float rz, float t, float E, float dr, int type, float scale) {
  p.r[0] = rx, p.r[1] = ry, p.r[2] = rz, p.r[3] = t;
  if (type < 0) {
    p.u = pfl.u;
    switch (type) {
      case -1: dr = 0, p.u.l.ka = -2; break;  // linear cone around n
      case -2: dr = 0; break;                 // 2-d gaussian around n
      case -3: dr = 0, p.u.l.ka = -1; break;  // isotropic
      case -4:
      default:
        dr = pfl.n[3];
        p.type = pfl.type;
        memcpy(p.r, pfl.r, sizeof p.r);
        memcpy(p.n, pfl.n, sizeof p.n);
    }
  } else {
    p.type = -type;
    p.u.p.beta = 1, p.u.p.tau = 0;
  }
  return scale * yield(E, dr, type) / det.ovr;
}

void addp(float rx, float ry, float rz, float t, float E, float dr, int type, float scale) {  // gpu/f2k.cxx:484-504
  addh(poidev(particle_setup(rx, ry, rz, t, E, dr, type, scale)));
}

// set p for a segment of n photons of a particle of velocity beta (gpu/f2k.cxx:507-518)
static void clst_setup(float rx, float ry, float rz, float t, float dr, float beta) {
  p.type = 0;
  p.r[3] = t, p.r[0] = rx, p.r[1] = ry, p.r[2] = rz;
  p.n[3] = dr, p.f = 1.f, p.u.p.a = 0, p.u.p.b = 0;
  p.u.p.beta = beta, p.u.p.tau = 0;
}

void addp_clst(float rx, float ry, float rz, float t, unsigned long long n, float dr, float beta) {
  clst_setup(rx, ry, rz, t, dr, beta);
  float qmas = 1.f;
  unsigned long long num = qmas > 0 && qmas < 1 ? bnldev(n, qmas) : n;
  addh(num);
}

// set p for a segment with light yield `light`, luminescence time tau and bare-track fraction
// fr (gpu/f2k.cxx:520-528); returns the Poisson mean of its photon count
static double mopo_setup(float rx, float ry, float rz, float t, float light, float length, float beta, float tau,
                         float fr) {
  p.type = 0;
  p.r[3] = t, p.r[0] = rx, p.r[1] = ry, p.r[2] = rz;
  p.n[3] = length, p.f = fr, p.u.p.a = 0, p.u.p.b = 0;
  p.u.p.beta = beta, p.u.p.tau = tau;
  return light * det.eff / det.ovr;
}

void addp_mopo(float rx, float ry, float rz, float t, float light, float length, float beta, float tau, float fr) {
  addh(poidev(mopo_setup(rx, ry, rz, t, light, length, beta, tau, fr)));
}

static void finc() {
  flne++;
  if ((int)(flne - flnb) < 0) {
    cerr << "Error: too many event segments" << endl;
    exit(3);
  }
}

void eout() {  // flush both the pending bunches and the launch in flight (gpu/f2k.cxx:536-541)
  output();
  output();
}

void efin() { hitz.clear(); }

void set_res(float t) { trz = t; }
void set_res(float t, float c) { trz = t, crz = c; }

void sett(float nx, float ny, float nz, std::pair<int, unsigned long long> id, int frame, double t0) {
  mcid ID;
  ID.first = id.first, ID.second = id.second, ID.frame = frame, ID.t0 = t0;
  flnz_lib.push_back(ID);
  finc();
  p.q = flne, p.n[0] = nx, p.n[1] = ny, p.n[2] = nz;
}

void flone(unsigned long long num) { addh(num * (det.eff / dppm) / det.ovr); }  // gpu/f2k.cxx:586-588

void flshift(float m[], float n[], float* r) {  // rotate (0,0,1) onto the DOM tilt m (gpu/f2k.cxx:590-628)
  if (std::isnan(n[2])) {
    float xi;
    sincosf(p.u.l.up, &n[2], &xi);
    n[0] *= xi, n[1] *= xi;
  }
  float q[3], o[3] = {0, 0, 1};
  q[0] = m[1] * o[2] - m[2] * o[1];
  q[1] = m[2] * o[0] - m[0] * o[2];
  q[2] = m[0] * o[1] - m[1] * o[0];
  float norm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  if (norm > 0) {
    float cs = 0;
    for (int i = 0; i < 3; i++) q[i] /= norm, cs += o[i] * m[i];
    float sn = sin(acos(cs));
    float R[3][3] = {{0}};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        R[i][j] = (i == j ? cs : sn * q[3 - i - j] * ((j - i + 3) % 3 == 1 ? 1 : -1)) + (1 - cs) * q[i] * q[j];
    for (int i = 0; i < 3; i++) {
      q[i] = 0;
      for (int j = 0; j < 3; j++) q[i] += R[i][j] * n[j];
    }
    for (int i = 0; i < 3; i++) n[i] = q[i];
    if (r != NULL) {
      for (int i = 0; i < 3; i++) {
        q[i] = 0;
        for (int j = 0; j < 3; j++) q[i] += R[i][j] * r[j];
      }
      for (int i = 0; i < 3; i++) r[i] = q[i];
    }
  }
}

const DOM& flset(int str, int dom) {  // configure a flasher or standard candle (gpu/f2k.cxx:630-759)
  p.u.l.fla = -1, p.u.l.ofla = -1;
  p.u.l.ka = 0, p.u.l.up = 0, p.u.l.fldr = -1;
  p.n[0] = 0, p.n[1] = 0, p.n[2] = 1, p.n[3] = 0;
  int type = 1;
  static DOM om;
  {
    float r[3] = {0, 0, 0};
    if (str < 0) type = 2, str = -str;
    if (str == 0) switch (dom) {  // standard candles
        case 1: type = 3, r[0] = 544.07, r[1] = 55.89, r[2] = 136.86; break;
        case 2: type = 4, r[0] = 11.87, r[1] = 179.19, r[2] = -205.64; break;
      }
    else
      for (int n = 0; n < det.t.gsize; n++)
        if (det.names[n].str == str && det.names[n].dom == dom) {
          p.u.l.fla = n;
          for (int m = 0; m < 3; m++) r[m] = det.oms[n].r[m];
          break;
        }
    cerr << "Flasher configured at " << r[0] << ", " << r[1] << ", " << r[2] << endl;
    p.type = type;
    p.r[0] = r[0], p.r[1] = r[1], p.r[2] = r[2], p.r[3] = 0;
    for (int m = 0; m < 3; m++) om.r[m] = r[m];
  }
  float fzcr = 0.f;
  if (const char* e = getenv("FZCR")) {
    fzcr = atof(e);
    cerr << "Adjusting flasher LED tilt by " << fzcr << " degrees" << endl;
  }
  switch (type) {
    case 1: p.u.l.up = (-0.2f + fzcr) * fcv; break;
    case 2: p.u.l.up = (48.1f + fzcr) * fcv; break;
    case 3: p.u.l.up = (90.0f - 41.13f) * fcv; break;
    case 4: p.u.l.up = (41.13f - 90.0f) * fcv; break;
  }
  if (const char* e = getenv("ELEV")) {
    float elev = atof(e);
    cerr << "Setting LED beam elevation angle to " << elev << " degrees" << endl;
    p.u.l.up = elev * fcv;
  }
  p.u.l.ofla = p.u.l.fla;
  if (const char* e = getenv("OFLA"))
    if (*e != 0 && atoi(e) == 0) {
      p.u.l.ofla = -1;
      cerr << "Flasher DOM will register photons!" << endl;
    }
  {
    const char* e = getenv("FLDR");
    p.u.l.fldr = e == nullptr ? -1 : atof(e);
    if (p.u.l.fldr >= 0) {
      float fold = int(p.u.l.fldr / 360);
      float dir = p.u.l.fldr - 360 * fold++;
      cerr << "Flasher LEDs are in a " << fold << "-fold pattern with LED #0 at " << dir << " degrees" << endl;
    }
  }
  float vx, vy;
  {
    float xi = p.u.l.fldr * fcv;
    sincosf(xi, &vy, &vx);
    sincosf(p.u.l.up, &p.n[2], &xi);
    p.n[0] = xi * vx, p.n[1] = xi * vy;
  }
  float fwid = 9.7f;
  if (const char* e = getenv("FWID")) {
    fwid = atof(e);
    cerr << "Setting flasher beam width to " << fwid << " degrees" << endl;
  }
  if (fwid > 999.f) {
    p.u.l.ka = fwid;
    if (p.u.l.fla >= 0 && p.u.l.fldr >= 0 && p.u.l.fldr < 360) {
      if (const char* e = getenv("FTLT")) {
        float m[3] = {0, 0, 1};
        float nx, ny, nz;
        int num = sscanf(e, "%f %f %f", &nx, &ny, &nz);
        if (num == 3) {
          cerr << "Applying tilt to flasher DOM of ";
          m[0] = nx, m[1] = ny, m[2] = nz;
        } else {
          ikey k;
          k.str = str, k.dom = dom;
          auto it = det.cx.find(k);
          if (it != det.cx.end()) {
            for (int i = 0; i < 3; i++) m[i] = it->second[i];
            cerr << "Setting flasher tilt from file cx.dat to ";
          } else
            cerr << "Leaving flasher tilt at default of ";
        }
        cerr << m[0] << ", " << m[1] << ", " << m[2] << endl;
        p.n[3] = 1;
        const float FLZ = 0.07735f, FLR = 0.119f - 0.008f;
        p.r[0] = FLR * vx, p.r[1] = FLR * vy, p.r[2] = FLZ;
        float n[3] = {p.n[0], p.n[1], p.n[2]};
        float r[3] = {p.r[0], p.r[1], p.r[2]};
        flshift(m, n, r);
        p.r[0] = r[0], p.r[1] = r[1], p.r[2] = r[2];
        p.n[0] = n[0], p.n[1] = n[1], p.n[2] = n[2];
      }
    }
  } else if (fwid > 0)
    p.u.l.ka = (fcv * fwid) * (fcv * fwid);
  else
    p.u.l.ka = fwid;
  pfl = p;
  return om;
}

const float* fldir() {
  static float dir[3] = {pfl.n[0], pfl.n[1], pfl.n[2]};
  return dir;
}

ppc_hip_bunch flasher_bunch(int str, int dom) {
  flset(str, dom);
  return p;
}

ppc_hip_bunch clst_bunch(float rx, float ry, float rz, float t, float nx, float ny, float nz, float dr, float beta) {
  p.n[0] = nx, p.n[1] = ny, p.n[2] = nz;
  clst_setup(rx, ry, rz, t, dr, beta);
  return p;
}

double mopo_bunch(float rx, float ry, float rz, float t, float nx, float ny, float nz, float light, float length,
                  float beta, float tau, float fr, ppc_hip_bunch& out) {
  p.n[0] = nx, p.n[1] = ny, p.n[2] = nz;
  double mean = mopo_setup(rx, ry, rz, t, light, length, beta, tau, fr);
  out = p;
  return mean;
}

double particle_bunch(float rx, float ry, float rz, float t, float E, float dr, int type, float nx, float ny,
                      float nz, ppc_hip_bunch& out) {
  p.n[0] = nx, p.n[1] = ny, p.n[2] = nz;
  double mean = particle_setup(rx, ry, rz, t, E, dr, type, 1.f);
  out = p;
  return mean;
}

// ------------------------------------------------------------ hit post-processing

static bool hit_less_full(const ppc_hip_hit& a, const ppc_hip_hit& b) {  // DTMN order (gpu/f2k.cxx:237-245)
  if (a.n != b.n) return (int)(a.n - b.n) < 0;
  if (a.i != b.i) return a.i < b.i;
  if (a.t != b.t) return a.t < b.t;
  return a.z < b.z;
}

// acceptance (RDE, angular sensitivity, PMT, cable shadow) and output of the raw hits (gpu/f2k.cxx:247-363)
// stable sort of a launch's accepted photons by hitz key: chunks sorted on their own
// threads, then adjacent chunks merged pairwise (both steps keep equal keys in order)
void sort_by_key(std::vector<std::pair<ihit, pout>>& b) {
  auto less = [](const std::pair<ihit, pout>& x, const std::pair<ihit, pout>& y) { return x.first < y.first; };
  const size_t n = b.size();
  const unsigned T = std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
  if (n < 65536 || T < 2) {
    std::stable_sort(b.begin(), b.end(), less);
    return;
  }
  std::vector<size_t> cut(T + 1);
  for (unsigned k = 0; k <= T; k++) cut[k] = n * k / T;
  std::vector<std::thread> th;
  for (unsigned k = 0; k < T; k++)
    th.emplace_back([&b, &cut, &less, k] { std::stable_sort(b.begin() + cut[k], b.begin() + cut[k + 1], less); });
  for (auto& t : th) t.join();
  for (unsigned w = 1; w < T; w *= 2) {
    th.clear();
    for (unsigned k = 0; k + w < T; k += 2 * w) {
      const size_t lo = cut[k], mid = cut[k + w], hi = cut[std::min(k + 2 * w, T)];
      th.emplace_back([&b, &less, lo, mid, hi] { std::inplace_merge(b.begin() + lo, b.begin() + mid, b.begin() + hi, less); });
    }
    for (auto& t : th) t.join();
  }
}

void print() {
  ppc_hip_hit* hs = be.hits.data();
  unsigned nh = be.hidx;
  const bool dev = be.dev_accept;
  std::vector<uint32_t> order;  // device acceptance: sort an index, the records stay in hit order
  if (dev) {
    order.resize(nh);
    for (unsigned i = 0; i < nh; i++) order[i] = i;
    if (deterministic())
      std::sort(order.begin(), order.end(), [hs](uint32_t a, uint32_t b) { return hit_less_full(hs[a], hs[b]); });
    else if ((int)(flnb - flnd) < 0)
      std::stable_sort(order.begin(), order.end(), [hs](uint32_t a, uint32_t b) { return (int)(hs[a].n - hs[b].n) < 0; });
  } else if (deterministic())
    std::sort(hs, hs + nh, hit_less_full);
  else if ((int)(flnb - flnd) < 0)
    std::stable_sort(hs, hs + nh, [](const ppc_hip_hit& a, const ppc_hip_hit& b) { return (int)(a.n - b.n) < 0; });

  // one accepted photon: a HIT line, or an entry of hitz (gpu/f2k.cxx:326-352), collected here
  // and put into the map in key order below
  std::vector<std::pair<ihit, pout>> batch;
  auto output = [&batch](const ppc_hip_hit& h, const DomInfo& n, int pmt, float wv, const float r[3],
                         const float d[3]) {
    if (cli_mode) {
      if (nextgen)
        printf("HIT %d %d_%d %f %f %f %f %f %f\n", n.str, n.dom, pmt, h.t, wv, h.pth, h.pph, h.dth, h.dph);
      else
        printf("HIT %d %d %f %f %f %f %f %f\n", n.str, n.dom, h.t, wv, h.pth, h.pph, h.dth, h.dph);
    } else {
      tmph.omkey = n;
      tmph.pmt = pmt;
      tmph.time = res(h.t, trz);
      tmph.dir = res(1 + cos(h.pth), crz) - 1;
      pout po;
      po.r[0] = r[0], po.r[1] = r[1], po.r[2] = r[2], po.r[3] = h.t;
      po.n[0] = d[0], po.n[1] = d[1], po.n[2] = d[2], po.n[3] = wv;
      batch.emplace_back(tmph, po);
    }
  };

  for (unsigned q = 0; q < nh; q++) {
    const unsigned idx = dev ? order[q] : q;
    ppc_hip_hit& h = hs[idx];
    if ((int)(h.n - flnb) < 0 || (int)(h.n - flnd) > 0) {
      cerr << "Internal Error: " << h.n << " is not between " << flnb << " " << flnd << " " << flne << endl;
      continue;
    }
    for (; (int)(flnb - h.n) < 0; flnb++) pop_segment();
    const DomInfo& n = det.names[h.i];
    if (dev) {
      const ppc_hip_accepted& a = be.acc[idx];
      if (a.pmt >= 0) {
        h.t += a.dt;
        output(h, n, a.pmt, a.wv, a.r, a.n);
      }
      continue;
    }
    bool flag = n.hv > 0;
    if (flag) {
      Acceptance& mt = det.accept[std::make_pair(n.omt, n.type)];
      flag = mt.rde > 0 && xrnd() < mt.rat[h.z] * n.rde / mt.rmax;
    }
    float nx = 0, ny = 0, nz = 0, rx = 0, ry = 0, rz = 0;
    if (flag) {
      float F = det.oms[h.i].F;
      float ptc = cosf(h.pth), pts = sinf(h.pth), ppc = cosf(h.pph), pps = sinf(h.pph);
      float dtc = F * cosf(h.dth), dts = F < 0 ? 1 : sinf(h.dth), dpc = cosf(h.dph), dps = sinf(h.dph);
      rx = dts * dpc, ry = dts * dps, rz = dtc;
      nx = pts * ppc, ny = pts * pps, nz = ptc;
    }
    int pmt = -1;
    if (flag) {
      auto it = det.types.find(n.omt);
      if (it != det.types.end()) {
        Vec3 dir, pos;
        dir[0] = nx, dir[1] = ny, dir[2] = nz;
        pos[0] = rx, pos[1] = ry, pos[2] = rz;
        pmt = it->second.pick_pmt(dir, pos, n.tilt, xrnd(), h.t, n.azi);
        flag = pmt >= 0;
      } else
        flag = false;
    }
    if (flag) {  // cable shadow
      float omr = det.oms[h.i].R;
      float rc = 0.023f, dc = omr + rc;
      float ph = n.azi;
      if (ph >= 0) {
        float drx = -omr * rx - dc * cos(fcv * ph);
        float dry = -omr * ry - dc * sin(fcv * ph);
        float a = nx * nx + ny * ny;
        if (a > 0) {
          float b = -2 * (nx * drx + ny * dry);
          float c = drx * drx + dry * dry - rc * rc;
          float D = b * b - 4 * a * c;
          if (D >= 0) {
            float h1 = (-b + sqrt(D)) / (2 * a);
            if (h1 > 0) flag = false;
          }
        }
      }
    }
    if (flag) {
      const WvBin& wb = det.wvs[h.z];
      float wv = wb.i + (wb.f - wb.i) * xrnd();
      float omr = det.oms[h.i].R;
      const float r[3] = {-omr * rx, -omr * ry, -omr * rz}, d[3] = {nx, ny, nz};
      output(h, n, pmt, wv, r, d);
    }
  }
  for (; (int)(flnb - flnd) < 0; flnb++) pop_segment();
  if (cli_mode) fflush(stdout);
  // hitz gets the same entries, in the same order within an entry, as one insertion per
  // photon would give; sorted by key, each insertion lands next to the previous one
  sort_by_key(batch);
  auto hint = batch.empty() ? hitz.end() : hitz.lower_bound(batch.front().first);
  for (const auto& e : batch) {
    auto it = hitz.emplace_hint(hint, std::piecewise_construct, std::forward_as_tuple(e.first), std::forward_as_tuple());
    it->second.push_back(e.second);
    hint = std::next(it);
  }
}

// ------------------------------------------------------------ command-line sources

void f2k() {  // particles from f2k "TR" lines on stdin (gpu/f2k.cxx:555-583)
  ini();
  string in;
  while (getline(std::cin, in)) {
    flnz_cli.push_back(in);
    finc();
    char name[32];
    int gens, igen, type = -1;
    float x, y, z, th, ph, l, E, t;
    const char* fmt = "TR %d %d %31s %f %f %f %f %f %f %f %f";
    if (sscanf(in.c_str(), fmt, &gens, &igen, name, &x, &y, &z, &th, &ph, &l, &E, &t) == 11) {
      th = fcv * (180 - th);
      ph = fcv * (ph < 180 ? ph + 180 : ph - 180);
      float costh = cosf(th), sinth = sinf(th), cosph = cosf(ph), sinph = sinf(ph);
      p.q = flne;
      p.n[0] = sinth * cosph, p.n[1] = sinth * sinph, p.n[2] = costh;
      if (0 == strcmp(name, "amu+") || 0 == strcmp(name, "amu-") || 0 == strcmp(name, "amu"))
        type = 0;
      else if (0 == strcmp(name, "delta") || 0 == strcmp(name, "brems") || 0 == strcmp(name, "epair") ||
               0 == strcmp(name, "e"))
        type = 11;
      else if (0 == strcmp(name, "e-"))
        type = 12;
      else if (0 == strcmp(name, "e+"))
        type = 13;
      else if (0 == strcmp(name, "munu") || 0 == strcmp(name, "hadr"))
        type = 101;
      else if (0 == strcmp(name, "ph"))
        addf(x, y, z, t, E);
      if (type >= 0) addp(x, y, z, t, E, l, type);
    }
  }
  eout();
  fin();
}

void flasher(int str, int dom, unsigned long long num, int itr) {  // gpu/f2k.cxx:768-783
  ini();
  flset(str, dom);
  for (int j = 0; j < std::max(1, itr); j++) {
    p.q = flne;
    flone(num);
    if (itr > 0) {
      flnz_cli.push_back("");
      finc();
    }
  }
  eout();
  fin();
}

// ------------------------------------------------------------ diagnostics

float host_zshift(const ppc_hip_tables& d, const float r[4], float& dz) {  // gpu/pro.cu:168-264
  const float* lp = det.lp.data();
  auto LP = [&](int i, int k, int c) { return lp[(i * LYRS + k) * 2 + c]; };
  if (d.tmod == 1) {
    float nr = d.lnx * r[0] + d.lny * r[1] - d.r0;
    int j = 1;
    for (; j < LMAX; j++)
      if (nr < d.lr[j] || j == d.lnum - 1) break;
    float z = (r[2] - d.lmin) * d.lrdz;
    int i = j - 1, k = std::min(std::max((int)floorf(z), 0), d.lpts);
    float dw = d.lr[j] - d.lr[i];
    float wx = (nr - d.lr[i]) / dw, wy = (d.lr[j] - nr) / dw;
    float qx = LP(j, k, 0) * wx + LP(i, k, 0) * wy;
    float qy = LP(j, k, 1) * wx + LP(i, k, 1) * wy;
    if (d.vthk != 0) dz /= 1 - d.lrdz * qy;
    return qx + qy * (z - k);
  } else if (d.tmod == 2) {
    float qx = d.cb[0] * r[0] + d.cb[2] * r[1];
    float qy = d.cb[1] * r[0] + d.cb[3] * r[1];
    qx = (qx - d.mmin[0]) / d.mstp[0];
    qy = (qy - d.mmin[1]) / d.mstp[1];
    float x = qx, y = qy, xy;
    int n0 = d.mnum[0] - 1, n1 = d.mnum[1] - 1, c0 = d.mcut[0] - 1, c1 = d.mcut[1] - 1;
    if (y < 0 && x - y > 0 && x <= c0) x = (x - y) * c0 / (c0 - y), y = 0;
    if (x < 0 && y - x >= 0 && y < c1) y = (y - x) * c1 / (c1 - x), x = 0;
    if (y > n1 && x >= n1 - c1 && y - x > n1 - n0)
      x = n1 - c1 + (n0 - (n1 - c1)) * (x - (n1 - c1)) / (y - (n1 - n0) - (n1 - c1)), y = n1;
    if (x > n0 && y > n0 - c0 && y - x <= n1 - n0)
      y = n0 - c0 + (n1 - (n0 - c0)) * (y - (n0 - c0)) / (x - (n0 - n1) - (n0 - c0)), x = n0;
    if (y - x > c1 && y >= c1 && x < n1 - c1) {
      xy = y - x - c1;
      xy = c1 + (x + y - c1 + xy) * (n1 - c1) / (n1 - c1 + xy);
      x = (xy - c1) / 2, y = (xy + c1) / 2;
    }
    if (x - y > c0 && x > c0 && y <= n0 - c0) {
      xy = x - y - c0;
      xy = c0 + (x + y - c0 + xy) * (n0 - c0) / (n0 - c0 + xy);
      x = (xy + c0) / 2, y = (xy - c0) / 2;
    }
    qx = x, qy = y;
    int nx = std::min(std::max((int)floorf(qx), 0), d.mnum[0]);
    int ny = std::min(std::max((int)floorf(qy), 0), d.mnum[1]);
    float fx = qx - nx, fy = qy - ny;
    float qrx, qry;
    int qnx, qny;
    if (fx > fy)
      qrx = fx, qry = fy, qnx = nx + 1, qny = ny;
    else
      qrx = fy, qry = fx, qnx = nx, qny = ny + 1;
    auto col = [&](int a, int b) { return a >= 0 && a < d.mnum[0] && b >= 0 && b < d.mnum[1] ? d.mcol[a * CTY + b] : -1; };
    int c0_ = col(nx, ny), c1_ = col(qnx, qny), c2_ = col(nx + 1, ny + 1);
    float w0 = 1 - qrx, w1 = qrx - qry, w2 = qry;
    float z = (r[2] - d.lmin) * d.lrdz;
    int k = std::min(std::max((int)floorf(z), 0), d.lpts);
    float ax = 0, ay = 0;
    if (c0_ >= 0) ax += LP(c0_, k, 0) * w0, ay += LP(c0_, k, 1) * w0;
    if (c1_ >= 0) ax += LP(c1_, k, 0) * w1, ay += LP(c1_, k, 1) * w1;
    if (c2_ >= 0) ax += LP(c2_, k, 0) * w2, ay += LP(c2_, k, 1) * w2;
    if (d.vthk != 0) dz /= 1 - d.lrdz * ay;
    return ax + ay * (z - k);
  }
  return 0;
}

void initialize(float enh) {  // gpu/ppc.cu:36
  det.build();
  det.eff *= enh;
}

}  // namespace xppc
