// ice.cpp -- builds the detector and ice tables from an ice directory.
//
// Restates ini::set() of the reference (gpu/ini.cxx:578-1700): same files, same
// environment variables, same arithmetic (in float where the reference uses
// float), same error exits.  The output goes into Detector: the device tables of
// ppc_hip.h (with the per-wavelength ice tables factorised into per-layer and
// per-wavelength terms) plus the dense tables and sensor data the host needs.
//
// Compiled with -ffast-math like the reference's host code (gpu/Makefile:7-10),
// so that the tables agree with the reference's bit for bit where the formulas
// are the same.
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "detector.h"

using std::cerr;
using std::endl;
using std::string;
using std::vector;

namespace xppc {

const float doma = FPI * OMR * OMR;  // DOM cross-section [m^2]
const float omav = 0.335822;         // average DOM angular sensitivity
const float dppm = 2450.08;          // photons per meter for a nominal DOM
const float fcv = FPI / 180.f;
const float zoff = 1948.07;

Detector det;
unsigned int sv = 0;

std::vector<OM> i3oms;
std::map<ikey, int> omts;
std::map<ikey, float> hvs;
std::map<ikey, std::pair<float, int> > rdes;
bool nextgen = false;

static void die(const string& msg) {
  cerr << msg << endl;
  exit(1);
}

// ------------------------------------------------------------ sensors

void SensorType::load_angsens(const string& file) {  // itype::add(string), ini.cxx:133-159
  def = true, mas = 1, ave = 0;
  area = 1, beta = 0.33f, Rr = OMR, Rz = OMR, cable = 0;
  std::ifstream in(file.c_str());
  if (in.fail()) die("Could not open file " + file);
  float aux;
  if (in >> aux) mas = aux;
  while (in >> aux) s.push_back(aux);
  if (s.empty()) s.push_back(1.f);
  if (mas > 0) {
    ave = s[0];
    for (unsigned int i = 2; i < s.size(); i += 2) ave += s[i] / i;
  } else
    ave = (1 - s[0]) / 2;
  if (ave > 0)
    cerr << "Loaded " << (s.size() + 1) << " angsens coefficients" << endl;
  else
    die("File " + file + " did not contain valid data");
  ave = omav;  // the reference keeps the nominal average for compatibility
  rde = (mas > 0 ? mas : 1.f) / (doma * ave);
}

void SensorType::add_pmt(float th, float ph) {  // ini.cxx:161-167
  float ct = std::cos(th * fcv), st = std::sin(th * fcv);
  float cp = std::cos(ph * fcv), sp = std::sin(ph * fcv);
  Vec3 d;
  d[0] = st * cp, d[1] = st * sp, d[2] = ct;
  dirs.push_back(d);
}

float SensorType::curve(float x) const {  // ini.cxx:169-188
  auto aS = [](float v) {
    float al = std::acos(v);
    return al - std::sin(2 * al) / 2;
  };
  if (beta < -1) return std::sqrt(1 - x * x);
  float sum = x > 0 ? x : 0;
  if (beta < 1) {
    float y = std::sqrt(1 - x * x) / beta;
    if (y > 1) {
      y = 1 / y;
      float c = 1 - beta * beta;
      sum += (aS(y) / c - aS(y * std::fabs(x) / std::sqrt(c)) * std::fabs(x)) / FPI;
    }
  }
  return sum;
}

float SensorType::xarea(float dot) const {  // ini.cxx:190-192
  return Rz > 0 ? FPI * Rr * std::sqrt(Rz * Rz - dot * dot * (Rz * Rz - Rr * Rr)) : -4 * Rr * Rz * std::sqrt(1 - dot * dot);
}

void SensorType::normalise(const vector<Vec3>& dirset) {  // fraq(), ini.cxx:194-208
  if (def) return;
  float fr_flat = 0, sum_ave = 0;
  for (const Vec3& i : dirset) {
    float tot = 0;
    for (const Vec3& j : dirs) tot += curve(i.dot(j));
    float fra = xarea(i[2]);
    fra = fra > 0 ? tot / fra : 0;
    if (fr_flat < fra) fr_flat = fra;
    sum_ave += tot;
  }
  sum_ave /= dirset.size();
  fx = fr_flat;
  rde = area * fr_flat / sum_ave;
}

int SensorType::pick_pmt(Vec3 dir, Vec3 pos, Vec3 tilt, float rnd, float& ht, float ph) const {  // ini.cxx:210-280
  if (beta == -2) {  // wavelength-shifting tube sensor
    const float cn = 0.299792458 / 1.46, absl = 3.0 / cn, att = 0.678, cut = 0.01, tau = 1.6, h = 0.76;
    const float a = h / (1 - std::exp(-h / att));
    float zz = Rr * pos[2];
    zz *= h / std::fabs(2 * Rz);
    int i = rnd - 0.5 < zz / (2 * a) ? 1 : 0;
    zz = i > 0 ? h / 2 - zz : h / 2 + zz;
    float d = (zz / h) * (1 - zz / (2 * a)) / (1 - h / (2 * a));
    zz = -att * std::log(1 - d * (1 - std::exp(-h / att)));
    if (zz < cut) zz = cut;
    float x = std::log(zz), dt;
    float s1 = std::exp(((0.0034112 * x + 0.041727) * x + 0.45550) * x - 1.9188);
    float g = std::exp(((0.0012658 * x + 0.017228) * x - 0.012750) * x + 0.20924);
    do {
      dt = s1 * std::sqrt(2 * g * std::pow(1 - xrnd(), 1 / (1 - g)) - 1);
    } while (dt > -absl * std::log(xrnd()));
    dt += zz / cn - tau * std::log(xrnd());
    ht += dt;
    return i;
  } else if (def) {
    bool flag;
    if (mas > 0) {
      float x = dir.dot(tilt), y = 1, sum = s[0];
      for (unsigned int i = 1; i < s.size(); i++) {
        y *= x;
        sum += s[i] * y;
      }
      flag = mas * rnd < sum;
    } else
      flag = pos.dot(tilt) > s[0];
    return flag ? 0 : -1;
  } else {
    if (ph >= 0) {  // rotate the photon by -ph instead of the PMTs
      ph -= cable;
      float cp = std::cos(fcv * ph), sp = std::sin(fcv * ph);
      float nx = dir[0] * cp + dir[1] * sp;
      float ny = dir[1] * cp - dir[0] * sp;
      dir[0] = nx, dir[1] = ny;
    }
    float crx = fx * xarea(dir[2]);
    unsigned int k = 0;
    float tot = 0;
    for (auto j = dirs.begin(); j != dirs.end(); ++j, k++) {
      tot += curve(-dir.dot(*j)) / crx;
      if (tot > rnd) break;
    }
    return k < dirs.size() ? (int)k : -1;
  }
}

// ------------------------------------------------------------ spectra

float Spectrum::index(float wv, float* ng) {  // phase and group refractive index, ini.cxx:301-305
  float np = 1.55749 - wv * (1.57988 - wv * (3.99993 - wv * (4.68271 - 2.09354 * wv)));
  if (ng != nullptr) *ng = np * (1 + 0.227106 - wv * (0.954648 - wv * (1.42568 - 0.711832 * wv)));
  return np;
}

float Spectrum::cherenkov(float wva) const {  // ini.cxx:307-316
  if (swv) return 1 / (wmax - wmin);
  const float c = FPI * 0.2 / 137.03599911;
  float wv = wva * 1.e-3, n = index(wv);
  return c * (1 - 1 / (n * n)) / (wv * wv);
}

// The reference scans past the end of the table when p exceeds its last
// (rounded) entry; here the scan stops at the last entry.
float Acceptance::inverse(float p) const {  // binf, ini.cxx:333-336
  const int last = (int)sum.size() - 1;
  int i, j = 0;
  while (j < last && sum[++j] <= 0) {
  }
  j--;
  while (j < last && sum[++j] < p) {
  }
  i = j - 1;
  return i + (sum[j] > sum[i] ? (p - sum[i]) / (sum[j] - sum[i]) : 0);
}

void Acceptance::sample(const Spectrum& sp, const vector<float>& qw, const vector<float>& qf, float blo,
                        float bhi) {  // read, ini.cxx:338-354
  dat.clear();
  int k = 0, n = qw.size();
  float wlo = qw[0] - (blo < 0 ? qw[1] - qw[0] : blo);
  float whi = qw[n - 1] + (bhi < 0 ? qw[n - 1] - qw[n - 2] : bhi);
  float w1 = wlo, w2 = qw[0], f1 = 0, f2 = qf[0];
  for (float w = sp.wmin + sp.bin / 2; w < sp.wmax; w += sp.bin) {
    while (k < n && qw[k] < w) {
      w1 = w2, f1 = f2;
      k++;
      if (k < n)
        w2 = qw[k], f2 = qf[k];
      else
        w2 = whi, f2 = 0;
    }
    dat.push_back(wlo < w && w < whi ? (f1 * (w2 - w) + f2 * (w - w1)) / (w2 - w1) : 0);
  }
}

// ------------------------------------------------------------ helpers

static string env_dir(const char* name, const string& fallback) {
  const char* e = getenv(name);
  return e != nullptr ? string(e) + "/" : fallback;
}

// line label of DOM n: strings 79-86 split at DOM 10 (DeepCore), high strings renumbered (ini.cxx:538-551)
static unsigned short line_label(const DomInfo& s) {
  static std::map<int, unsigned short> overflow;
  static unsigned short next_num = 86 + 10 + 1;
  if (s.str > 86) {
    auto it = overflow.find(s.str);
    if (it == overflow.end()) {
      if (next_num >= 0x8000) die("Number of string labels exceeds capacity of 32768");
      overflow.insert(std::make_pair(s.str, next_num++));
    }
    return overflow[s.str];
  }
  return (unsigned short)((s.str > 78 && s.dom > 10) ? s.str + 10 : s.str);
}

// tabulated tilt file: rows of depth followed by npts values (ini.cxx:1212-1243, 1272-1303)
static int read_tilt_rows(const string& file, int npts, vector<float>& ds, vector<vector<float> >& lp) {
  std::ifstream in(file.c_str());
  if (in.fail()) return -1;
  lp.assign(npts, vector<float>());
  vector<float> pts(npts);
  string line;
  while (getline(in, line)) {
    std::stringstream str(line);
    float depth;
    while (str >> depth) {
      int i = 0;
      while (str >> pts[i++])
        if (i >= npts) break;
      if (i != npts) break;
      ds.push_back(depth);
      for (i = 0; i < npts; i++) lp[i].push_back(pts[i]);
    }
  }
  return (int)ds.size();
}

// ------------------------------------------------------------ build

void Detector::build() {
  *this = Detector();  // a second initialisation starts from scratch
  ppc_hip_tables& d = t;

  string ppcdir("");
  {
    const char* e = getenv("PPCTABLESDIR");
    if (e != nullptr)
      ppcdir = string(e) + "/";
    else {
      e = getenv("I3_SRC");
      if (e != nullptr) ppcdir = string(e) + "/ppc/resources/ice/";
    }
    cerr << "Configuring ppc in \"" << ppcdir << "\"" << endl;
  }
  string omdir = env_dir("NEXTGENDIR", ppcdir);
  cerr << "Configuring nextgen sensors in \"" << omdir << "\"" << endl;
  string icedir = env_dir("ICEMODELDIR", ppcdir);
  cerr << "Configuring icemodel in \"" << icedir << "\"" << endl;
  string tiltdir = env_dir("TILTMODELDIR", icedir);
  cerr << "Configuring tiltmodel in \"" << tiltdir << "\"" << endl;
  string holeice;
  {
    const char* e = getenv("PPCHOLEICE");
    holeice = e != nullptr ? string(e) : ppcdir + "as.dat";
    cerr << "Configuring holeice from \"" << holeice << "\"" << endl;
  }

  float dk1, dk2, dkz;  // scattering anisotropy of cfg.txt

  {  // cfg.txt (ini.cxx:629-729)
    std::ifstream in((icedir + "cfg.txt").c_str());
    if (in.fail()) die("Could not open file cfg.txt");
    string line;
    float aux;
    vector<float> v;
    while (getline(in, line))
      if (sscanf(line.c_str(), "%f", &aux) == 1) v.push_back(aux);
    if (const char* ov = getenv("OVSZ")) {
      float ovsz = atof(ov);
      if (v.size() >= 1) v[0] = ovsz;
      cerr << "Using oversize as set with OVSZ=" << ovsz << endl;
    }
    if (v.size() < 4) die("File cfg.txt did not contain valid data");

    int xR = lroundf(v[0]);
    d.xR = xR;
    ovr *= xR * xR;
    eff = v[1], d.sf = v[2], d.g = v[3];
    d.gr = (1 - d.g) / (1 + d.g);
    cerr << "Configured: xR=" << xR << " eff=" << eff << " sf=" << d.sf << " g=" << d.g << endl;

    if (v.size() < 12)
      d.SF = d.sf, d.G = d.g, d.GR = d.gr;
    else
      d.SF = v[10], d.G = v[11], d.GR = (1 - d.G) / (1 + d.G);

    if (v.size() >= 10) {
      float xH = v[7], hS = v[8], hA = v[9];
      d.hr = OMR * xH;
      d.hr2 = d.hr * d.hr;
      d.hs = 1 / (hS * (1 - d.G)), d.ha = 1 / hA;
      if (xH > 0) cerr << "With hole ice: xH=" << xH << " sca=" << hS << " (" << d.SF << "," << d.G << ") abs=" << hA << endl;
    } else
      d.hr = 0, d.hr2 = 0, d.hs = 0, d.ha = 0;

    if (v.size() >= 7) {
      const float thx = v[4];
      d.azx = std::cos(fcv * thx), d.azy = std::sin(fcv * thx);
      dk1 = std::exp(v[5]);
      dk2 = std::exp(v[6]);
      dkz = 1 / (dk1 * dk2);
      cerr << "Ice anisotropy is k(" << thx << ")=" << dk1 << "," << dk2 << "," << dkz << endl;
    } else
      dk1 = 1, dk2 = 1, dkz = 1, d.azx = 1, d.azy = 0;

    if (v.size() >= 15) {  // absorption anisotropy, renormalised for an axially symmetric case
      d.k1 = std::exp(v[12]);
      d.k2 = std::exp(v[13]);
      d.kz = std::exp(v[14]);
      cerr << "New Ice anisotropy is " << d.k1 << "," << d.k2 << "," << d.kz << endl;
      if (d.k1 >= d.k2 && d.k2 == d.kz) {
        float r = d.k1 / d.k2;
        float s = std::sqrt(r * r - 1);
        r = (s > XXX ? std::log(r + s) / s : 1 + s / 2) / d.k2;
        d.k1 *= r, d.k2 *= r, d.kz *= r;
        cerr << "Renorm. NI anisotropy " << d.k1 << "," << d.k2 << "," << d.kz << endl;
      } else
        cerr << "Warning: taking absorption anisotropy as given (not renormalizing)!" << endl;
    } else
      d.k1 = 1, d.k2 = 1, d.kz = 1;

    if (v.size() >= 16) {
      d.fr = v[15];
      cerr << "Ice absorption anisotropy scaling is " << d.fr << endl;
    } else
      d.fr = 1;

    if (v.size() >= 28) {  // birefringence
      for (int i = 0; i < 12; i++) d.bfr[i] = v[16 + i];
      {
        const char* A = getenv("BFRA");
        float bfra = A == nullptr ? 1.0 : atof(A);
        const char* B = getenv("BFRB");
        float bfrb = B == nullptr ? 1.0 : atof(B);
        if (A != nullptr || B != nullptr) cerr << "Setting BFRA=" << bfra << " BFRB=" << bfrb << endl;
        for (int i = 0; i < 12; i += 4) d.bfr[i] *= i < 8 ? std::sqrt(bfra) : bfra * bfrb;
      }
      {  // effective scattering of the diffusion pattern, midpoint rule over cos
        float step = 0.01, sum = 0;
        for (float x = step / 2; x < 1; x += step) {
          float y = std::sqrt(1 - x * x);
          float sx = std::max(0.f, d.bfr[0] * std::exp(-d.bfr[1] * std::pow(std::atan(d.bfr[3] * y), d.bfr[2])));
          float sy = std::max(0.f, d.bfr[4] * std::exp(-d.bfr[5] * std::pow(std::atan(d.bfr[7] * y), d.bfr[6])));
          float mx = std::max(0.f, d.bfr[8] * std::atan(d.bfr[11] * y * x) * std::exp(-d.bfr[9] * y + d.bfr[10] * x));
          sum += sx * sx + sy * sy + mx * mx;
        }
        sum *= step / 2;
        d.sum = sum;
      }
      cerr << "Initialized BFR diffusion patterns; s_eff=" << d.sum << " m^-1" << endl;
    } else {
      d.sum = 0;
      for (int i = 0; i < 12; i++) d.bfr[i] = i % 4 < 1 ? 0 : 1;
    }
  }

  {  // nextgen sensor types (ini.cxx:731-761)
    std::ifstream in((omdir + "om.conf").c_str());
    if (!in.fail()) {
      string line;
      while (getline(in, line)) {
        int m;
        unsigned int n;
        SensorType st;
        float th, ph, other;
        int read = sscanf(line.c_str(), "%*s %d %f %f %f %f %d %f %f %f", &m, &st.area, &st.beta, &st.Rr, &st.Rz, &n,
                          &th, &ph, &other);
        st.cable = read >= 9 ? other : 0;
        if (read >= 8) {
          st.add_pmt(th, ph);
          for (unsigned int i = 1; i < n && getline(in, line); i++)
            if (2 == sscanf(line.c_str(), "%f %f %f", &th, &ph, &other)) st.add_pmt(th, ph);
          if (st.dirs.size() == n) types.insert(std::make_pair(m, st));
        }
      }
    }
    if (!types.empty()) {
      std::ifstream di((omdir + "om.dirs").c_str());
      if (!di.fail()) {
        string line;
        while (getline(di, line)) {
          Vec3 v;
          if (3 == sscanf(line.c_str(), "%*d %f %f %f", &v[0], &v[1], &v[2])) ico.push_back(v);
        }
        cerr << "Configured " << ico.size() << " uniform directions" << endl;
      }
      if (ico.size() < 1) die("Error: could not initialize an array of directions");
      nextgen = true;
      for (auto& j : types) {
        j.second.normalise(ico);
        cerr << " OM Type " << j.first << " with " << j.second.dirs.size() << " PMTs added (" << j.second.rde << ")"
             << endl;
      }
    }
  }
  if (types.find(-1) == types.end()) types[-1].load_angsens(holeice);

  {  // MWC multipliers (ini.cxx:767-799)
    std::ifstream in((ppcdir + "rnd.txt").c_str());
    if (in.fail()) die("Could not open file rnd.txt");
    string line;
    while (getline(in, line)) {
      std::stringstream str(line);
      unsigned int a;
      if (str >> a) rm.push_back(a);
    }
    if (rm.size() < 1) die("File rnd.txt did not contain valid data");
    cerr << "Loaded " << rm.size() << " random multipliers" << endl;
    const char* dt = getenv("DTMN");
    if (dt == nullptr || atoi(dt) == 0) {  // time-of-day seed unless deterministic
      timeval tv;
      gettimeofday(&tv, NULL);
      sv = 1000000 * (unsigned long long)tv.tv_sec + tv.tv_usec;
    }
  }

  {  // geometry (ini.cxx:801-818)
    std::ifstream in((ppcdir + "geo-f2k").c_str());
    if (!in.fail()) {
      if (!i3oms.empty()) {
        i3oms.clear();
        cerr << "Warning: overwriting existing geometry!" << endl;
      }
      OM om;
      string mbid;
      unsigned long long omid;
      while (in >> mbid >> std::hex >> omid >> std::dec >> om.r[0] >> om.r[1] >> om.r[2] >> om.str >> om.dom) {
        om.R = OMR, om.F = 1;
        om.r[2] += zoff;
        i3oms.push_back(om);
      }
    }
  }
  {
    const char* e = getenv("HIFL");
    float hifl = e == nullptr ? 0.f : atof(e);
    if (e != nullptr) cerr << "Setting HIFL=" << hifl << " ns" << endl;
    d.hifl = hifl;
  }
  {  // GECO: move one DOM and/or change the hole ice (ini.cxx:826-855)
    std::map<ikey, std::pair<float, float> > geco;
    if (const char* e = getenv("GECO")) {
      ikey om;
      float x, y, xH, hS;
      int num = sscanf(e, "%d %d %f %f %f %f", &om.str, &om.dom, &x, &y, &xH, &hS);
      if (num >= 4) {
        geco[om] = std::make_pair(OMR * x, OMR * y);
        cerr << "Applying geometry correction to DOM " << om.str << "," << om.dom << " of " << x << "," << y
             << " x OMR (" << (100 * OMR) << " cm)" << endl;
      }
      if (num >= 5) {
        d.hr = OMR * xH;
        d.hr2 = d.hr * d.hr;
        cerr << "Updating hole ice parameters: xH=" << xH << endl;
      }
      if (num >= 6) {
        d.hs = 1 / (hS * (1 - d.G));
        cerr << "Updating hole ice parameters: sca=" << hS << endl;
      }
    }
    for (OM& o : i3oms) {
      auto j = geco.find(o);
      if (j != geco.end()) o.r[0] += j->second.first, o.r[1] += j->second.second;
    }
  }
  {
    std::ifstream in((ppcdir + "str-f2k").c_str());
    if (!in.fail()) {
      cerr << "Using fixed positions of hole ice columns from str-f2k!" << endl;
      int s;
      float x, y;
      while (in >> s >> x >> y) strs[s] = std::make_pair(x, y);
    }
  }
  {
    std::ifstream in((ppcdir + "eff-f2k").c_str());
    if (!in.fail()) {
      if (!rdes.empty()) {
        rdes.clear();
        cerr << "Warning: overwriting existing RDE table!" << endl;
      }
      int typ;
      ikey om;
      float e;
      string line;
      while (getline(in, line)) {
        int num = sscanf(line.c_str(), "%d %d %f %d", &om.str, &om.dom, &e, &typ);
        if (num < 4) typ = 0;
        if (num >= 3)
          if (om.isinice()) rdes.insert(std::make_pair(om, std::make_pair(e, typ)));
      }
    }
  }
  {
    std::ifstream in((omdir + "om.map").c_str());
    if (!in.fail()) {
      if (!omts.empty()) {
        omts.clear();
        cerr << "Warning: overwriting existing OM type table!" << endl;
      }
      ikey om;
      int m;
      while (in >> om.str >> om.dom >> m)
        if (om.isinice()) omts.insert(std::make_pair(om, m));
    }
  }
  {
    std::ifstream in((ppcdir + "hvs-f2k").c_str());
    if (!in.fail()) {
      if (!hvs.empty()) {
        hvs.clear();
        cerr << "Warning: overwriting existing HV table!" << endl;
      }
      ikey om;
      float hv;
      while (in >> om.str >> om.dom >> hv)
        if (om.isinice()) hvs.insert(std::make_pair(om, hv));
    }
  }
  {
    std::ifstream in((ppcdir + "cx.dat").c_str());
    if (!in.fail()) {
      ikey om;
      Vec3 dir;
      float r;
      while (in >> om.str >> om.dom >> dir[0] >> dir[1] >> dir[2] >> r) cx[om] = dir;
      if (cx.size() > 0)
        cerr << "Loaded " << cx.size() << " DOM orientations" << endl;
      else
        die("File cx.dat did not contain valid data");
    }
  }
  {
    std::ifstream in((ppcdir + "dx.dat").c_str());
    if (!in.fail()) {
      ikey om;
      float dir, r;
      while (in >> om.str >> om.dom >> dir >> r) {
        while (dir < 0) dir += 360.f;
        dx[om] = dir;
      }
      if (dx.size() > 0)
        cerr << "Loaded " << dx.size() << " cable positions" << endl;
      else
        die("File dx.dat did not contain valid data");
    }
  }

  float Rr = 0, Rz = 0;
  {  // in-ice DOMs, sorted by (string, dom) (ini.cxx:945-1017)
    vector<ppc_hip_dom> list;
    vector<DomInfo> info;
    int nhqe = 0;
    std::sort(i3oms.begin(), i3oms.end());
    for (OM& o : i3oms) {
      if (!o.isinice()) continue;
      ikey om(o);
      int m, ty;
      float r, h;
      DomInfo nm;
      float azi = -1.f;
      if (omts.empty())
        m = -1;
      else {
        auto j = omts.find(om);
        m = j == omts.end() ? -1 : j->second;
        if (types.find(m) == types.end()) m = -1;
        auto it = types.find(m);
        if (it != types.end()) o.R = it->second.Rr, o.F = it->second.Rz / it->second.Rr;
      }
      Rr = std::max(Rr, o.R), Rz = std::max(Rz, o.R * std::fabs(o.F));
      ppc_hip_dom dm;
      dm.R = o.R, dm.F = o.F, dm.r[0] = o.r[0], dm.r[1] = o.r[1], dm.r[2] = o.r[2];
      list.push_back(dm);
      {
        auto j = rdes.find(om);
        if (j != rdes.end()) {
          nhqe++;
          r = j->second.first;
          ty = j->second.second;
        } else
          r = 1, ty = 0;
        accept[std::make_pair(m, ty)].add_om(r);
      }
      if (hvs.empty())
        h = 1200;
      else {
        auto j = hvs.find(om);
        h = j == hvs.end() ? 0 : j->second;
      }
      {
        auto ci = cx.find(om);
        if (ci != cx.end()) nm.tilt = ci->second;
        auto di = dx.find(om);
        if (di != dx.end()) azi = di->second;
      }
      nm.str = om.str, nm.dom = om.dom, nm.omt = m, nm.type = ty, nm.rde = r, nm.hv = h, nm.azi = azi;
      info.push_back(nm);
    }
    if (nhqe > 0) cerr << "Loaded " << nhqe << " RDE coefficients" << endl;
    for (auto& i : accept) cerr << " Found " << i.second.oms << " OMs of type " << i.first.first << "," << i.first.second << endl;
    int gsize = list.size();
    if (gsize > MAXGEO) {
      cerr << "Error: too many OMs (" << gsize << "), truncating to " << MAXGEO << endl;
      gsize = MAXGEO;
    }
    oms.assign(list.begin(), list.begin() + gsize);
    names.assign(info.begin(), info.begin() + gsize);
    d.gsize = gsize;
  }

  Rr *= d.xR, Rz *= d.xR;

  // strings ("lines") of the cell grid (ini.cxx:1021-1081)
  std::map<unsigned short, short> num;
  vector<ppc_hip_line> lines;
  {
    std::map<unsigned short, float> low;
    std::map<unsigned short, ppc_hip_line> acc;
    for (int n = 0; n < d.gsize; n++) {
      unsigned short str = line_label(names[n]);
      ppc_hip_line& s = acc[str];
      const ppc_hip_dom& om = oms[n];
      if (num.find(str) == num.end()) {
        low[str] = om.r[2];
        s.h = om.r[2];
        s.n = n;
      } else {
        if (low[str] > om.r[2]) low[str] = om.r[2];
        if (s.h < om.r[2]) s.h = om.r[2];
        if (s.n > n) s.n = n;
      }
      num[str]++;
      s.x += om.r[0], s.y += om.r[1];
    }
    if (acc.size() > (size_t)NSTR) die("Number of strings exceeds capacity of 240");
    for (auto& i : num) {
      unsigned short str = i.first, n = i.second;
      ppc_hip_line& s = acc[str];
      float dz = s.h - low[str];
      if (n > 1 && dz <= 0) {
        cerr << "Cannot estimate the spacing along string " << (int)str << endl;
        exit(1);
      }
      s.x /= n, s.y /= n, s.r = 0;
      s.d = n > 1 ? (n - 1) / dz : 0;
      s.dl = 0, s.dh = 0;
      auto j = strs.find(str);
      if (j != strs.end()) s.x = j->second.first, s.y = j->second.second;
    }
    for (int n = 0; n < d.gsize; n++) {
      unsigned short str = line_label(names[n]);
      ppc_hip_line& s = acc[str];
      const ppc_hip_dom& om = oms[n];
      float ddx = s.x - om.r[0], ddy = s.y - om.r[1];
      float dr = ddx * ddx + ddy * ddy;
      if (dr > s.r) s.r = dr;
      if (s.d > 0) {
        float dz = om.r[2] - (s.h + ((short)s.n - n) / s.d);
        if (s.dl > dz) s.dl = dz;
        if (s.dh < dz) s.dh = dz;
      }
    }
    d.rx = 0;
    int n = 0;
    for (auto i = num.begin(); i != num.end(); ++i, n++) {
      ppc_hip_line& s = acc[i->first];
      s.max = i->second - 1;
      i->second = n;
      s.r = Rr + std::sqrt(s.r);
      if (d.hr > s.r) s.r = d.hr;
      if (d.rx < s.r) d.rx = s.r;
      s.dl -= Rz, s.dh += Rz;
      lines.push_back(s);
    }
  }
  sc = lines;

  float sin12 = 0;
  {  // skewed cell coordinates along 9.3 and 129.3 deg (ini.cxx:1083-1101)
    float bv[2][2];
    bv[0][0] = std::cos(fcv * DIR1);
    bv[0][1] = std::sin(fcv * DIR1);
    bv[1][0] = std::cos(fcv * DIR2);
    bv[1][1] = std::sin(fcv * DIR2);
    float det_ = bv[0][0] * bv[1][1] - bv[0][1] * bv[1][0];
    d.cb[0] = bv[1][1] / det_;
    d.cb[1] = -bv[0][1] / det_;
    d.cb[2] = -bv[1][0] / det_;
    d.cb[3] = bv[0][0] / det_;
    for (int i = 0; i < 2; i++) sin12 += bv[0][i] * bv[1][i];
    sin12 = std::sqrt(1 - sin12 * sin12);
    d.rx /= sin12;
  }
  auto ctr = [&](const ppc_hip_line& s, int m) { return d.cb[0 * 2 + m] * s.x + d.cb[1 * 2 + m] * s.y; };

  std::map<unsigned short, int> cells[CX][CY];
  {  // cell grid (ini.cxx:1103-1157)
    float cl[2] = {0, 0}, ch[2] = {0, 0}, crst[2];
    int n = 0;
    for (auto i = num.begin(); i != num.end(); ++i, n++) {
      const ppc_hip_line& s = sc[i->second];
      for (int m = 0; m < 2; m++) {
        if (n == 0 || ctr(s, m) < cl[m]) cl[m] = ctr(s, m);
        if (n == 0 || ctr(s, m) > ch[m]) ch[m] = ctr(s, m);
      }
    }
    d.cn[0] = CX;
    d.cn[1] = CY;
    for (int m = 0; m < 2; m++) {
      float diff = ch[m] - cl[m];
      d.cn[m] = std::min(d.cn[m], 1 + 2 * (int)lroundf(diff / 125));
      if (d.cn[m] <= 1) {
        ch[m] = cl[m] = (cl[m] + ch[m]) / 2;
        crst[m] = 1 / (d.rx * (2 + XXX) + diff);
      } else {
        float s = Rr * (d.cn[m] - 1);
        if (diff < 2 * s) {
          cerr << "Warning: tight string packing in direction " << (m < 1 ? "x" : "y") << endl;
          float ave = (cl[m] + ch[m]) / 2;
          cl[m] = ave - s;
          ch[m] = ave + s;
          diff = 2 * s;
        }
        crst[m] = (d.cn[m] - 1) / diff;
      }
    }
    bool flag = true;
    for (auto i = num.begin(); i != num.end(); ++i) {
      const ppc_hip_line& s = sc[i->second];
      int k[2];
      for (int m = 0; m < 2; m++) {
        k[m] = lroundf((ctr(s, m) - cl[m]) * crst[m]);
        if (k[m] < 0 || k[m] >= d.cn[m]) die("Error in cell initialization");
        float d1 = std::fabs(ctr(s, m) - (cl[m] + (k[m] - 0.5f) / crst[m]));
        float d2 = std::fabs(ctr(s, m) - (cl[m] + (k[m] + 0.5f) / crst[m]));
        float dd = std::min(d1, d2) * sin12 - s.r;
        if (dd < 0) {
          flag = false;
          cerr << "Warning: string " << (int)i->first << " too close to cell boundary" << endl;
        }
      }
      cells[k[0]][k[1]][i->first]++;
    }
    if (flag) d.rx = 0;
    for (int m = 0; m < 2; m++) d.cl[m] = cl[m], d.crst[m] = crst[m];
  }
  {  // cell -> string lists (ini.cxx:1159-1174)
    is.assign(CX * CY, 0);
    ls.assign(NSTR, 0);
    unsigned int pos = 0;
    for (int i = 0; i < d.cn[0]; i++)
      for (int j = 0; j < d.cn[1]; j++) {
        auto& c = cells[i][j];
        if (c.size() > 0) {
          is[i * CY + j] = pos;
          for (auto& s : c) {
            if (pos == (unsigned)NSTR) die("Number of string cells exceeds capacity of 240");
            ls[pos++] = num[s.first];
          }
          ls[pos - 1] |= 0x8000;
        } else
          is[i * CY + j] = 0x8000;
      }
  }
  cerr << "Loaded " << d.gsize << " DOMs (" << d.cn[0] << "x" << d.cn[1] << ")" << endl;

  d.tmod = 0;
  lp.assign(LMAX * LYRS * 2, 0.f);
  auto store_tilt = [&](const vector<float>& ds, const vector<vector<float> >& pts) {
    int size = ds.size();
    for (int i = 0; i < d.lnum; i++)
      for (int j = 0; j < size - 1; j++) {
        lp[(i * LYRS + j) * 2 + 0] = pts[i][size - 1 - j];
        lp[(i * LYRS + j) * 2 + 1] = pts[i][size - 2 - j] - pts[i][size - 1 - j];
      }
    d.lpts = size - 2;
    if (size < 2)
      d.lnum = 0;
    else {
      float lmin = ds[0], lmax = ds[size - 1];
      d.lmin = zoff - lmax;
      d.lrdz = (size - 1) / (lmax - lmin);
    }
  };
  {  // 2-d tilt map (ini.cxx:1180-1245)
    std::ifstream in((tiltdir + "tilt.set").c_str());
    if (!in.fail()) {
      float mdir[2];
      int n = 0;
      string line;
      while (getline(in, line)) {
        int tot = sscanf(line.c_str(), "%f %f %f %d %d", &mdir[n], &d.mmin[n], &d.mstp[n], &d.mnum[n], &d.mcut[n]);
        if (tot == 5) n++;
        if (n >= 2) break;
      }
      if (n < 2) die("File tilt.set found, but is corrupt");
      if (mdir[0] != DIR1 || mdir[1] != DIR2) die("Unsupported tilt grid configuration");
      int size = 0;
      for (int j = 0; j < d.mnum[1]; j++)
        for (int i = 0; i < d.mnum[0]; i++)
          if (i < j + d.mcut[0] && j < i + d.mcut[1]) size++;
      if (size > LMAX) die("File tilt.set defines too many dust maps");
      if (d.mnum[0] > CTX || d.mnum[1] > CTY) die("Error: tilt map conifguration exceeds buffer");
      if (d.mstp[0] <= 0 || d.mstp[1] <= 0) die("Tilt map does not use increasing range order");
      for (int m = 0, j = 0; j < d.mnum[1]; j++)
        for (int i = 0; i < d.mnum[0]; i++) d.mcol[i * CTY + j] = (i < j + d.mcut[0] && j < i + d.mcut[1]) ? m++ : -1;
      vector<float> ds;
      vector<vector<float> > pts;
      if (read_tilt_rows(tiltdir + "tilt.map", size, ds, pts) >= 0) {
        d.lnum = size;
        int rows = ds.size();
        if (rows - 1 > LYRS) die("File tilt.map defines too many map points");
        for (int i = 1; i < rows; i++)
          if (ds[i] < ds[i - 1]) die("Tilt map does not use increasing depth order");
        store_tilt(ds, pts);
        if (d.lnum > 0) {
          d.tmod = 2;
          cerr << "Loaded " << d.lnum << "x" << rows << " 2d dust layer points" << endl;
        }
      }
    }
  }
  if (d.lnum <= 0) {  // 1-d tilt along 225 deg (ini.cxx:1247-1305)
    d.lnum = 0;
    l0 = 0, d.r0 = 0;
    const float thx = 225;
    d.lnx = std::cos(fcv * thx), d.lny = std::sin(fcv * thx);
    std::ifstream in((tiltdir + "tilt.par").c_str());
    if (!in.fail()) {
      int str;
      float aux;
      vector<float> lr;
      while (in >> str >> aux) {
        if (aux == 0) l0 = str;
        lr.push_back(aux);
      }
      int size = lr.size();
      if (size > LMAX) die("File tilt.par defines too many dust maps");
      for (int i = 1; i < size; i++)
        if (lr[i] < lr[i - 1]) die("Tilt map does not use increasing range order");
      for (int i = 0; i < size; i++) d.lr[i] = lr[i];
      vector<float> ds;
      vector<vector<float> > pts;
      if (read_tilt_rows(tiltdir + "tilt.dat", size, ds, pts) >= 0) {
        d.lnum = size;
        int rows = ds.size();
        if (rows - 1 > LYRS) die("File tilt.dat defines too many map points");
        for (int i = 1; i < rows; i++)
          if (ds[i] < ds[i - 1]) die("Tilt map does not use increasing depth order");
        store_tilt(ds, pts);
        if (d.lnum > 0) {
          d.tmod = 1;
          cerr << "Loaded " << d.lnum << "x" << rows << " dust layer points" << endl;
        }
      }
    }
  }
  {  // r0: mean projection of the zero-distance string (ini.cxx:1307-1319)
    int nr = 0;
    float r0 = 0;
    for (int n = 0; n < d.gsize; n++)
      if (names[n].str == l0) {
        nr++;
        r0 += d.lnx * oms[n].r[0] + d.lny * oms[n].r[1];
      }
    if (nr > 0) d.r0 = r0 / nr;
  }
  {
    const char* e = getenv("VTHK");
    d.vthk = e == nullptr ? 0 : atoi(e);
    cerr << "Ice layer thickness: " << (d.vthk == 0 ? "UNIFORM" : d.vthk == 1 ? "TILT-CORRECTED" : "UNKNOWN") << endl;
  }

  // ---- ice and wavelength tables (ini.cxx:1338-1696) ----
  {
    vector<float> wx, wy, wz;
    bool flag = true, rdef = false;
    const char* WFLA = getenv("WFLA");
    float wfla = WFLA == nullptr ? 0 : atof(WFLA);
    if (wfla > 0) {  // single LED wavelength, +-10 nm smoothstep
      for (float x = 0.f; x < 1.f + XXX; x += 0.1f) {
        wx.push_back(3 * x * x - 2 * x * x * x);
        wy.push_back(wfla + 10.f * (2 * x - 1));
      }
      cerr << "Using single wavelength=" << wfla << " [nm]" << endl;
      spec.single(wfla);
    } else {
      string flwl("dat");
      if (const char* e = getenv("FLWL")) {
        flwl = string(e);
        cerr << "wv.dat file extension is modified with FLWL=" << flwl << endl;
      }
      flwl = "wv." + flwl;
      std::ifstream in((ppcdir + flwl).c_str());
      if (in.fail()) die("Could not open file " + flwl);
      int n = 0;
      float xa, ya, xo = 0, yo = 0;
      while (in >> xa >> ya) {
        if ((xa < 0 || 1 < xa) || (n == 0 && xa != 0) || (n > 0 && (xa <= xo || ya <= yo))) {
          flag = false;
          break;
        }
        wx.push_back(xa);
        wy.push_back(ya);
        xo = xa, yo = ya;
        n++;
      }
      if (xo != 1 || wx.size() < 2) flag = false;
      if (flag)
        cerr << "Loaded " << wx.size() << " wavelenth points" << endl;
      else
        die("File " + flwl + " did not contain valid data");
    }
    if (flag) {
      vector<float> qx, qy;
      std::ifstream in((ppcdir + "wv.rde").c_str());
      if (!in.fail()) {
        int n = 0;
        bool ok = true;
        float xa, ya, yo = 0;
        while (in >> ya >> xa) {
          if (xa < 0 || (n > 0 && ya <= yo)) {
            ok = false;
            break;
          }
          qx.push_back(xa);
          qy.push_back(ya);
          yo = ya;
          n++;
        }
        if (qx.size() < 2) ok = false;
        if (ok)
          cerr << "Loaded " << qx.size() << " RDE coefficients" << endl;
        else
          die("File wv.rde did not contain valid data");
        rdef = ok;
      }
      if (rdef) {
        int k = 0, n = qy.size();
        for (float w : wy) {
          float r;
          for (; k < n; k++)
            if (qy[k] > w) break;
          if (k == 0)
            r = qx[0];
          else if (k == n)
            r = qx[n - 1];
          else {
            r = ((w - qy[k - 1]) * qx[k] + (qy[k] - w) * qx[k - 1]) / (qy[k] - qy[k - 1]);
            k--;
          }
          wz.push_back(r);
        }
      }
      vector<float> qw, qf, qz;
      float qo = 0, wo = wy[0];
      unsigned int n = wx.size();
      for (unsigned int i = 1; i < n; i++) {
        float qv = wx[i], wv_ = wy[i];
        float bin = wv_ - wo;
        float wva = wv_ - bin / 2;
        float val = dppm * doma * omav * (qv - qo) / bin;
        qw.push_back(wva);
        qf.push_back(val);
        if (rdef) qz.push_back(val * (wz[i] + wz[i - 1]) / 2);
        qo = qv, wo = wv_;
      }
      float blo = wy[1] - wy[0], bhi = wy[n - 1] - wy[n - 2];
      accept[std::make_pair(-1, 0)].sample(spec, qw, qf, blo, bhi);
      if (rdef) accept[std::make_pair(-1, 1)].sample(spec, qw, qz, blo, bhi);
    }
  }
  for (auto& i : accept) {  // per-sensor-type acceptance files (ini.cxx:1440-1469)
    const std::pair<int, int>& m = i.first;
    if (m.first < 0) continue;
    std::stringstream str;
    str << omdir << "om.wv_" << m.first << "." << m.second;
    const string file = str.str();
    bool flag = false;
    vector<float> qw, qf;
    std::ifstream in(file.c_str());
    if (!in.fail()) {
      int n = 0;
      flag = true;
      float w, f, u = 0;
      while (in >> w >> f) {
        if (f < 0 || (n > 0 && w <= u)) {
          flag = false;
          break;
        }
        qf.push_back(f * spec.cherenkov(w));
        qw.push_back(w);
        u = w;
        n++;
      }
      if (qf.size() < 2) flag = false;
      if (flag)
        cerr << "Loaded " << qf.size() << " wv points from file " << file << endl;
      else
        die("File " + file + " did not contain valid data");
    }
    if (flag)
      i.second.sample(spec, qw, qf);
    else
      cerr << "Warning: file " << file << " was not read. Omitting hits in this OM type!" << endl;
  }
  {  // envelope of all acceptances; its integral is the photon yield per meter (ini.cxx:1471-1492)
    float sum = 0;
    int k = 0;
    for (float w = spec.wmin + spec.bin / 2; w < spec.wmax; w += spec.bin, k++) {
      float val = 0;
      for (auto& i : accept) {
        Acceptance& mt = i.second;
        if (!mt.dat.empty()) {
          float ri = mt.dat[k] * types[i.first.first].rde * mt.rmax * spec.bin;
          mt.rde += ri;
          mt.sum.push_back(mt.rde);
          mt.dat[k] = ri;
          if (val < ri) val = ri;
        }
      }
      sum += val;
      env.dat.push_back(val);
      env.sum.push_back(sum);
    }
    env.rde = sum;
    eff *= sum;
    cerr << "Calculated value of ppm=" << eff << endl;
    for (int i = 0; i < k; i++) env.sum[i] /= sum;
  }

  float wv0 = 400;
  float A, B, Dc, E, a, k;
  float Ae, Be, De, Ee, ae, ke;
  vector<float> dp, be, ba, td, k1, k2, ra, rb;
  {  // icemodel.par (ini.cxx:1499-1513)
    std::ifstream in((icedir + "icemodel.par").c_str());
    if (in.fail()) die("File icemodel.par was not found");
    bool flag = true, fail = false;
    if (flag) flag = (bool)(in >> a >> ae);
    if (flag) flag = (bool)(in >> k >> ke);
    if (flag) flag = (bool)(in >> A >> Ae);
    if (flag) flag = (bool)(in >> B >> Be);
    fail = !flag;
    if (flag) flag = (bool)(in >> Dc >> De);
    Dc = flag ? Dc / std::pow(wv0, k) : 1;
    if (flag) flag = (bool)(in >> E >> Ee);
    E = flag ? E / std::pow(wv0, k) : 0;
    if (fail) die("File icemodel.par found, but is corrupt");
  }
  int size = 0;
  {  // icemodel.dat (ini.cxx:1515-1541)
    std::ifstream in((icedir + "icemodel.dat").c_str());
    if (in.fail()) die("File icemodel.dat was not found");
    float dpa, bea, baa, tda, k1a, k2a, bfra, bfrb;
    string line;
    while (getline(in, line)) {
      int n = sscanf(line.c_str(), "%f %f %f %f %f %f %f %f", &dpa, &bea, &baa, &tda, &k1a, &k2a, &bfra, &bfrb);
      if (n >= 4) {
        dp.push_back(dpa);
        be.push_back(bea);
        ba.push_back(baa);
        td.push_back(tda);
        k1.push_back(n >= 5 ? std::exp(k1a) : dk1);
        k2.push_back(n >= 6 ? std::exp(k2a) : dk2);
        ra.push_back(n >= 7 ? bfra : 1);
        rb.push_back(n >= 8 ? bfrb : 1);
        size++;
      }
    }
    if (size < 1) die("File icemodel.dat found, but is corrupt");
  }
  (void)dkz;
  float dh = size > 1 ? (dp[size - 1] - dp[0]) / (size - 1) : 100.f;
  if (dh <= 0) die("Ice table does not use increasing depth spacing");
  for (int i = 1; i < size; i++)
    if (std::fabs(dp[i] - dp[i - 1] - dh) > dh * XXX) die("Ice table does not use uniform depth spacing");
  cerr << "Loaded " << size << " ice layers" << endl;
  if (size > MAXLYS) {
    cerr << "Error: too many layers (" << size << "), truncating to " << MAXLYS << endl;
    size = MAXLYS;
  }
  d.size = size;
  d.dh = dh;
  d.rdh = 1 / dh;
  d.hmin = zoff - dp[size - 1];

  float bble = 0, bblz = 0, bbly = 0;
  {
    std::ifstream in((icedir + "icemodel.bbl").c_str());
    if (!in.fail()) {
      if (!(in >> bble >> bblz >> bbly)) {
        cerr << "File icemodel.bbl found, but is corrupt" << endl;
        bble = 0;
      } else
        cerr << "Air bubble parameters: " << bble << " " << bblz << " " << bbly << endl;
    }
  }
  az.assign(size, ppc_hip_aniz{});
  for (int i = 0; i < size; i++) {  // layer 0 is the deepest (ini.cxx:1579-1583)
    int j = size - 1 - i;
    az[i].k1 = k1[j], az[i].k2 = k2[j];
    az[i].ra = ra[j], az[i].rb = ra[j] * rb[j];
  }
  float arf = 0, srw = 0, srf = 1;
  {
    int version = 0;
    std::ifstream in((icedir + "version").c_str());
    if (!in.fail()) {
      if (!(in >> version)) {
        cerr << "File version found, but is corrupt" << endl;
        version = 0;
      } else
        cerr << "Adjusting defaults for ice model version = " << version << endl;
    }
    switch (version) {
      case 10: arf = 0, srw = 0, srf = 0; break;
      case 20: arf = 1, srw = 1, srf = 0; break;
    }
  }
  if (const char* e = getenv("ABSM")) {
    cerr << "Ice absorption table: ";
    switch (atoi(e)) {
      case 0: arf = 0, cerr << "MieDUST"; break;
      case 1: arf = 1, cerr << "FULL400"; break;
      default: cerr << "UNKNOWN";
    }
    cerr << endl;
  }
  if (const char* e = getenv("BFRM")) {
    cerr << "BFR scattering subtraction method: ";
    switch (atoi(e)) {
      case 0: srw = 0, srf = 1, cerr << "DEFAULT"; break;
      case 1: srw = 1, srf = 0, cerr << "CORRECT"; break;
      case 2: srw = 0, srf = 0, cerr << "NO SUB."; break;
      default: cerr << "UNKNOWN";
    }
    cerr << endl;
  }

  // per-layer factors of the ice tables: sca = s0 + s1*la, abs = p*lk + t*abl
  lay.assign(size, ppc_hip_layer{});
  vector<float> bbl(size);
  for (int i = 0; i < size; i++) {
    int j = size - 1 - i;
    bbl[i] = bble > 0 ? dp[j] < bblz && dp[j] < bbly ? bble * (1 - dp[j] / bblz) * (1 - dp[j] / bbly) : 0 : 0;
    float tdf = 1 + 0.01 * td[j];
    lay[i].s0 = (bbl[i] - ra[j] * d.sum * srf) / (1 - d.g);
    lay[i].s1 = (be[j] - ra[j] * d.sum * srw) / (1 - d.g);
    lay[i].p = Dc * ba[j] + E;
    lay[i].t = tdf;
  }
  const float AB0 = A * std::exp(-B / wv0);
  for (int i = 0; i < size; i++) lay[i].p -= arf * AB0 * lay[i].t;
  par_a.resize(size), par_b.resize(size), dpda.assign(size, Dc), ds1db.assign(size, 1 / (1 - d.g));
  for (int i = 0; i < size; i++) par_a[i] = ba[size - 1 - i], par_b[i] = be[size - 1 - i];

  wv.assign(WNUM, ppc_hip_wv{});
  abs_tab.assign(WNUM * MAXLYS, 0.f);
  sca_tab.assign(WNUM * MAXLYS, 0.f);
  for (auto& i : accept) i.second.rat.clear();
  for (int n = 0; n < WNUM; n++) {  // equal-probability wavelength bins (ini.cxx:1651-1696)
    float wvi = env.inverse((n + 0.f) / WNUM);
    int mi = std::min(std::max((int)std::floor(wvi), 0), spec.num - 1);
    float wvf = env.inverse((n + 1.f) / WNUM);
    int mf = std::min(std::max((int)std::floor(wvf), 0), spec.num - 1);
    float wva = spec.wav(env.inverse((n + 0.5f) / WNUM));
    wvs[n].w = wva, wvs[n].i = spec.wav(wvi), wvs[n].f = spec.wav(wvf);
    for (auto& i : accept) {
      Acceptance& mt = i.second;
      if (mt.rde > 0) {
        const vector<float>& wy = mt.dat;
        float r;
        if (mf == mi)
          r = wy[mi] * (wvf - wvi);
        else {
          r = (mi + 1 - wvi) * wy[mi] + (wvf - mf) * wy[mf];
          for (int k2i = mi + 1; k2i < mf; k2i++) r += wy[k2i];
        }
        r /= env.rde / WNUM;
        mt.rat.push_back(r);
      }
    }
    float l_a = std::pow(wva / wv0, -a);
    float l_k = std::pow(wva / wv0, -k);
    float ABl = A * std::exp(-B / wva);
    for (int i = 0; i < size; i++) {
      int j = size - 1 - i;
      float sca = (bbl[i] + be[j] * l_a - ra[j] * d.sum * (srf + srw * l_a)) / (1 - d.g);
      float abs = (Dc * ba[j] + E) * l_k + (ABl - arf * AB0 * l_k) * (1 + 0.01 * td[j]);
      if (sca > 0 && abs > 0)
        sca_tab[n * MAXLYS + i] = sca, abs_tab[n * MAXLYS + i] = abs;
      else
        die("Invalid value of ice parameter, cannot proceed");
    }
    float ng, np = Spectrum::index(wva * 1.e-3, &ng);
    float c = 0.299792458;
    d.ocv = 1 / c;
    wv[n].wvl = n;
    wv[n].ocm = ng / c;
    wv[n].coschr = 1 / np;
    wv[n].sinchr = std::sqrt(1 - wv[n].coschr * wv[n].coschr);
    wv[n].la = l_a, wv[n].lk = l_k, wv[n].abl = ABl;
  }
  cerr << endl;
  refresh_pointers();
}

void Detector::refresh_pointers() {
  t.lay = lay.data();
  t.az = az.data();
  t.wv = wv.data();
  t.is = is.data();
  t.ls = ls.data();
  t.nsc = (int)sc.size();
  t.sc = sc.data();
  t.lp = lp.data();
  t.oms = oms.data();
  t.rsize = (int)rm.size();
  t.rm = rm.data();
}

}  // namespace xppc
