// detector.h -- host-side state of libxppc: the tables built from an ice
// directory (the reference keeps these in the globals d, z, q and m of
// gpu/ini.cxx) and the helpers the hit post-processing needs.
#ifndef XPPC_DETECTOR_H
#define XPPC_DETECTOR_H

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "ppc.h"
#include "ppc_hip.h"

namespace xppc {

// capacities and constants of the reference (gpu/ini.cxx:1-36, 553)
constexpr int LMAX = PPC_HIP_LMAX;
constexpr int LYRS = PPC_HIP_LYRS;
constexpr int CTX = PPC_HIP_CTX, CTY = PPC_HIP_CTY;
constexpr int CX = PPC_HIP_CX, CY = PPC_HIP_CY;
constexpr int NSTR = PPC_HIP_NSTR;
constexpr int WNUM = PPC_HIP_WNUM;
constexpr int MAXLYS = PPC_HIP_MAXLYS;
constexpr int MAXGEO = 16384;
constexpr int MAXRND = 131072;
constexpr float XXX = 1.e-5f;
constexpr float FPI = 3.141592653589793f;
constexpr float OMR = 0.16510f;
constexpr float DIR1 = 9.3f, DIR2 = 129.3f;
extern const float doma, omav, dppm, fcv, zoff;

// 3-vector with the reference's default (0,0,1) (gpu/ini.cxx:67-84)
struct Vec3 {
  float x[3] = {0, 0, 1};
  float& operator[](int i) { return x[i]; }
  float operator[](int i) const { return x[i]; }
  float dot(const Vec3& q) const {
    float s = 0;
    for (int i = 0; i < 3; i++) s += x[i] * q.x[i];
    return s;
  }
};

// sensor type: angular sensitivity and PMT layout (gpu/ini.cxx:122-281)
struct SensorType {
  float area = 0, beta = 0, rde = 0, fx = 0, Rr = 0, Rz = 0, cable = 0;
  std::vector<Vec3> dirs;
  bool def = false;
  float ave = 0, mas = 0;
  std::vector<float> s;

  void load_angsens(const std::string& file);
  void add_pmt(float th, float ph);
  float curve(float x) const;
  float xarea(float dot) const;
  void normalise(const std::vector<Vec3>& ico);
  int pick_pmt(Vec3 dir, Vec3 pos, Vec3 tilt, float rnd, float& ht, float ph) const;
};

// wavelength grid of the acceptance tables (gpu/ini.cxx:283-317)
struct Spectrum {
  bool swv = false;
  int num = 450;
  float bin, wmin = 250.f, wmax = 700.f;
  Spectrum() { bin = (wmax - wmin) / num; }
  void single(float wva) {
    swv = true, num = 10, wmin = wva - 5.f, wmax = wva + 5.f;
    bin = (wmax - wmin) / num;
  }
  float wav(float i) const { return wmin + bin * i; }
  static float index(float wv, float* ng = nullptr);
  float cherenkov(float wva) const;
};

// wavelength acceptance of one (sensor type, PMT type) class (gpu/ini.cxx:319-355)
struct Acceptance {
  float rde = 0;
  std::vector<float> dat, sum, rat;
  int oms = 0;
  float rmax = 0;
  void add_om(float r) {
    if (r > rmax) rmax = r;
    oms++;
  }
  float inverse(float p) const;
  void sample(const Spectrum& sp, const std::vector<float>& qw, const std::vector<float>& qf, float blo = -1.f,
              float bhi = -1.f);
};

// per-DOM metadata used by the hit post-processing (gpu/ini.cxx:86-98)
struct DomInfo : ikey {
  int omt = -1, type = 0;
  float rde = 1, hv = 0, azi = -1;
  Vec3 tilt;
};

// the wavelength draw within a bin (gpu/ini.cxx:523-529)
struct WvBin {
  float w = 0, i = 0, f = 0;
};

struct Detector {
  // ---- what the device needs (ppc_hip_tables points into these) ----
  ppc_hip_tables t{};
  std::vector<ppc_hip_layer> lay;
  std::vector<ppc_hip_aniz> az;
  std::vector<ppc_hip_wv> wv;
  std::vector<uint16_t> is, ls;
  std::vector<ppc_hip_line> sc;
  std::vector<float> lp;  // [LMAX][LYRS][2]
  std::vector<ppc_hip_dom> oms;
  std::vector<uint32_t> rm;

  // ---- dense per-wavelength ice tables as the reference stores them ----
  std::vector<float> abs_tab, sca_tab;  // [WNUM][MAXLYS]
  // per layer (layer 0 deepest): the icemodel.dat absorption and scattering parameters
  // adust(400) and be(400), and the derivatives of the layer factors p and s1 by them
  std::vector<float> par_a, par_b, dpda, ds1db;
  int l0 = 0;

  // ---- host post-processing ----
  std::vector<DomInfo> names;
  std::map<int, SensorType> types;
  std::map<std::pair<int, int>, Acceptance> accept;
  Acceptance env;
  Spectrum spec;
  std::vector<Vec3> ico;
  WvBin wvs[WNUM];
  float eff = 0;
  unsigned int ovr = 1;
  std::map<ikey, Vec3> cx;
  std::map<ikey, float> dx;
  std::map<int, std::pair<float, float> > strs;

  void build();  // ini::set() (gpu/ini.cxx:578-1700)
  void refresh_pointers();
};

extern Detector det;
extern unsigned int sv;  // host RNG state, also the MWC seed offset (gpu/ini.cxx:554)

}  // namespace xppc

#endif
