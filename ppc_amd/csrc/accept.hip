// accept.hip -- the reference's hit post-processing (gpu/f2k.cxx:247-363 with
// itype::getPMT, gpu/ini.cxx:169-280) as a device epilogue: one lane per raw hit
// decides the relative DOM efficiency, the PMT (angular sensitivity, the PMT
// layout of a multi-PMT sensor, or the end of a wavelength-shifting tube with its
// transit time), the cable shadow and the photon's wavelength within its bin, and
// writes the accepted photon's impact point, direction and added time.
//
// Random numbers: the reference draws them from one host stream in hit order;
// here every draw is a counter-based uniform of (seed, hit, draw) (SplitMix64
// finaliser, Steele, Lea and Flood, OOPSLA 2014), so the device result is
// statistically the reference's, not the same numbers.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "kparams.h"
#include "ppc_hip.h"

namespace ppc {

namespace {

constexpr float kPI = 3.141592653589793f;
constexpr float kFCV = kPI / 180.f;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform in (0, 1): draw number `draw` of the hit with key `key`
__device__ __forceinline__ float urand(uint64_t key, uint32_t draw) {
  const uint64_t z = mix64(key + (draw + 1u) * 0x9E3779B97F4A7C15ull);
  return ((float)(uint32_t)(z >> 40) + 0.5f) * (1.f / 16777216.f);
}

// PMT acceptance curve of a multi-PMT sensor (gpu/ini.cxx:169-188)
__device__ float curve(const ppc_hip_sensor& S, float x) {
  if (S.beta < -1) return sqrtf(1 - x * x);
  float sum = x > 0 ? x : 0;
  if (S.beta < 1) {
    float y = sqrtf(1 - x * x) / S.beta;
    if (y > 1) {
      y = 1 / y;
      const float c = 1 - S.beta * S.beta;
      auto aS = [](float v) {
        const float al = acosf(v);
        return al - sinf(2 * al) / 2;
      };
      sum += (aS(y) / c - aS(y * fabsf(x) / sqrtf(c)) * fabsf(x)) / kPI;
    }
  }
  return sum;
}

// projected area of the sensor seen along a direction with z component dot (gpu/ini.cxx:190-192)
__device__ float xarea(const ppc_hip_sensor& S, float dot) {
  return S.Rz > 0 ? kPI * S.Rr * sqrtf(S.Rz * S.Rz - dot * dot * (S.Rz * S.Rz - S.Rr * S.Rr))
                  : -4 * S.Rr * S.Rz * sqrtf(1 - dot * dot);
}

// Wavelength-shifting tube sensor (gpu/ini.cxx:211-243, after the reference's model of a
// painted quartz tube read out at both ends): pz is the axial component of the hit's
// position vector (the impact point is minus the tube radius times it).  The photon is shifted at a depth into the tube and
// guided to one of the two ends (the return value, chosen with rnd against the attenuation
// along the tube); dt is the transit time: the guided path at the speed of light in quartz,
// a delay of the parameterised spread of guided paths (drawn by rejection against absorption
// in the quartz) and the paint's exponential re-emission.  Draws 3.. of the hit; the rejection
// loop stops after kTries pairs (a try is rejected with probability at most 0.21, so all
// of them with probability below 1e-43).
constexpr int kTries = 64;
__device__ int wls_tube(const ppc_hip_sensor& S, float pz, float rnd, uint64_t key, float& dt) {
  constexpr float cq = 0.299792458f / 1.46f;  // speed of light in quartz [m/ns]
  constexpr float tq = 3.0f / cq;              // absorption in quartz [ns]
  constexpr float att = 0.678f;                // attenuation length along the tube [m]
  constexpr float zmin = 0.01f;                // lower end of the spread's parameterisation [m]
  constexpr float tau = 1.6f;                  // re-emission time of the paint [ns]
  constexpr float len = 0.76f;                 // tube length [m]
  const float ea = expf(-len / att), a = len / (1 - ea);
  float z = S.Rr * pz * (len / fabsf(2 * S.Rz));
  const int end = rnd - 0.5f < z / (2 * a) ? 1 : 0;
  z = end ? len / 2 - z : len / 2 + z;
  const float d = (z / len) * (1 - z / (2 * a)) / (1 - len / (2 * a));
  z = -att * logf(1 - d * (1 - ea));
  if (z < zmin) z = zmin;
  const float x = logf(z);
  const float s = expf(((0.0034112f * x + 0.041727f) * x + 0.45550f) * x - 1.9188f);
  const float g = expf(((0.0012658f * x + 0.017228f) * x - 0.012750f) * x + 0.20924f);
  float t = 0;
  for (int k = 0; k < kTries; k++) {
    t = s * sqrtf(2 * g * powf(1 - urand(key, 4 + 2 * k), 1 / (1 - g)) - 1);
    if (!(t > -tq * logf(urand(key, 5 + 2 * k)))) break;
  }
  dt = t + z / cq - tau * logf(urand(key, 3));
  return end;
}

}  // namespace

// The random numbers of a hit depend on the seed and on the hit itself (its photon, DOM and
// time), not on its slot in the hit buffer, so a deterministic launch is post-processed
// the same way whatever order its hits were stored in.
__global__ void accept_hits(AcceptParams A, const ppc_hip_hit* __restrict__ hits, const uint32_t* __restrict__ photon,
                            uint32_t n, uint64_t seed, ppc_hip_accepted* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ppc_hip_hit h = hits[i];
  const uint64_t key = mix64(seed ^ mix64(((uint64_t)photon[i] << 32 | h.i) ^ ((uint64_t)__float_as_uint(h.t) << 17)));
  ppc_hip_accepted o;
  o.pmt = -1, o.wv = 0;
  o.r[0] = o.r[1] = o.r[2] = 0, o.n[0] = o.n[1] = o.n[2] = 0, o.dt = 0;
  if ((int32_t)h.i < 0 || (int32_t)h.i >= A.gsize || h.z >= PPC_HIP_WNUM) {
    out[i] = o;
    return;
  }
  const ppc_hip_dom_accept d = A.doms[h.i];
  bool flag = d.sensor >= 0 && urand(key, 0) < A.rat[(size_t)d.acc * PPC_HIP_WNUM + h.z] * d.rde;
  if (flag) {
    float pts, ptc, pps, ppc, dts, dtc, dps, dpc;
    sincosf(h.pth, &pts, &ptc);
    sincosf(h.pph, &pps, &ppc);
    sincosf(h.dth, &dts, &dtc);
    sincosf(h.dph, &dps, &dpc);
    dtc *= d.F;
    if (d.F < 0) dts = 1;
    const float rx = dts * dpc, ry = dts * dps, rz = dtc;
    float nx = pts * ppc, ny = pts * pps, nz = ptc;
    const ppc_hip_sensor& S = A.sensors[d.sensor];
    const float rnd = urand(key, 1);
    int pmt = -1;
    float dt = 0;
    if (S.beta == -2)  // wavelength-shifting tube: the end that records the photon
      pmt = wls_tube(S, rz, rnd, key, dt);
    else if (S.def) {  // one PMT: angular sensitivity along the DOM axis
      if (S.mas > 0) {
        const float x = nx * d.tilt[0] + ny * d.tilt[1] + nz * d.tilt[2];
        float y = 1, sum = S.s[0];
        for (int k = 1; k < S.ns; k++) y *= x, sum += S.s[k] * y;
        pmt = S.mas * rnd < sum ? 0 : -1;
      } else
        pmt = rx * d.tilt[0] + ry * d.tilt[1] + rz * d.tilt[2] > S.s[0] ? 0 : -1;
    } else {  // PMTs of a multi-PMT sensor, the photon rotated by minus the cable azimuth
      float dx = nx, dy = ny;
      if (d.azi >= 0) {
        float sp, cp;
        sincosf(kFCV * (d.azi - S.cable), &sp, &cp);
        dx = nx * cp + ny * sp, dy = ny * cp - nx * sp;
      }
      const float crx = S.fx * xarea(S, nz);
      float tot = 0;
      int k = 0;
      for (; k < S.npmt; k++) {
        const float* q = A.pmt_dirs + 3 * (S.pmt0 + k);
        tot += curve(S, -(dx * q[0] + dy * q[1] + nz * q[2])) / crx;
        if (tot > rnd) break;
      }
      pmt = k < S.npmt ? k : -1;
    }
    flag = pmt >= 0;
    if (flag && d.azi >= 0) {  // cable shadow (gpu/f2k.cxx:301-322)
      const float omr = d.R, rc = 0.023f, dc = omr + rc;
      float sp, cp;
      sincosf(kFCV * d.azi, &sp, &cp);
      const float drx = -omr * rx - dc * cp, dry = -omr * ry - dc * sp;
      const float a = nx * nx + ny * ny;
      if (a > 0) {
        const float b = -2 * (nx * drx + ny * dry);
        const float c = drx * drx + dry * dry - rc * rc;
        const float D = b * b - 4 * a * c;
        if (D >= 0 && (-b + sqrtf(D)) / (2 * a) > 0) flag = false;
      }
    }
    if (flag) {
      const float lo = A.wvs[2 * h.z], hi = A.wvs[2 * h.z + 1];
      o.pmt = pmt;
      o.wv = lo + (hi - lo) * urand(key, 2);
      o.r[0] = -d.R * rx, o.r[1] = -d.R * ry, o.r[2] = -d.R * rz;
      o.n[0] = nx, o.n[1] = ny, o.n[2] = nz;
      o.dt = dt;
    }
  }
  out[i] = o;
}

hipError_t launch_accept(const AcceptParams& A, const ppc_hip_hit* hits, const uint32_t* photon, uint32_t n,
                         uint64_t seed, ppc_hip_accepted* out, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(accept_hits, dim3((n + 255) / 256), dim3(256), 0, stream, A, hits, photon, n, seed, out);
  return hipGetLastError();
}

}  // namespace ppc
