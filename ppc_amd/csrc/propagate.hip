// propagate.hip -- photon propagation kernel for CDNA4 (gfx950).
//
// One persistent launch: every lane owns one multiply-with-carry stream and
// propagates photons one after another.  Photons are handed out per wave: a wave
// claims a chunk of photon indices from the global counter with one atomic, and
// when at least `refill` of its 64 lanes are idle it gives them the next indices
// of its chunk; each idle lane emits its photon (source sampling) before the wave
// steps again (lanes still idle step along on a state that changes nothing).  One
// loop iteration is one transport step of the reference kernel
// (gpu/pro.cu:486-889): absorption anisotropy,
// layer walk to the next scattering/absorption point, DOM search through the
// cell grid, hit recording, birefringence, and scattering.  Hits are compacted
// per wave with ballot + prefix count into the HBM hit buffer.
//
// Physics and the order in which each photon consumes its MWC stream follow
// gpu/pro.cu:287-904 step by step and draw for draw, so that a photon replayed
// from its recorded stream state on the CPU gives the same path (FP32 arithmetic
// is rearranged where that only changes rounding).  The per-wavelength ice tables
// are evaluated from per-layer factors held in LDS (sca = s0 + s1*la,
// abs = p*lk + t*abl), instead of a 512 x layers table.
#include <hip/hip_runtime.h>

#include "kparams.h"

namespace ppc {

#define FPI 3.141592653589793f
constexpr float XXX = 1.e-5f;
constexpr float OMR = 0.16510f;

// ---------------------------------------------------------------- math
// Hardware transcendental instructions: v_log/v_exp are base 2, v_sin/v_cos
// take revolutions; all are accurate to about 1 ulp on their domain here.
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * 0.69314718056f; }
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504089f); }
__device__ __forceinline__ float fpow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * frcp(b); }
__device__ __forceinline__ float sq(float x) { return x * x; }
__device__ __forceinline__ int rn(float x) { return (int)__builtin_rintf(x); }
__device__ __forceinline__ int rd(float x) { return (int)__builtin_floorf(x); }
__device__ __forceinline__ int ru(float x) { return (int)__builtin_ceilf(x); }

// number of set bits of the wave mask m below this lane (v_mbcnt: no per-lane 64-bit mask)
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// element i of a table in global memory: a global (not flat) load whose base stays in
// scalar registers and whose byte offset is one 32-bit vector register (tables here are
// far below 4 GiB); a flat load would also hold up the next wait for LDS results
// (read as a plain float vector: the HIP vector types have no constructor from a global
// address-space reference)
typedef __attribute__((address_space(1))) const char gbyte;
template <class T>
__device__ __forceinline__ T ldg(const T* base, uint32_t i) {
  typedef __attribute__((ext_vector_type(sizeof(T) / 4))) float raw;
  const raw v = *(const __attribute__((address_space(1))) raw*)((gbyte*)base + i * (uint32_t)sizeof(T));
  T out;
  __builtin_memcpy(&out, &v, sizeof(T));
  return out;
}

// ---------------------------------------------------------------- RNG
// SplitMix64 finaliser (Steele, Lea and Flood, OOPSLA 2014): the start state of photon i
// when photons are seeded one by one
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Multiply-with-carry: t = a*x + c (one v_mad_u64_u32), x = lo(t), c = hi(t);
// uniform in (0,1] from the top 23 bits of x (gpu/pro.cu:47-57).  x and c leave the
// multiply as two 32-bit values (the empty asm): kept as one 64-bit register pair they
// are spilled and reloaded as a pair, which the gfx950 register allocator has placed at
// odd register numbers (an illegal encoding that fails the build).
struct Mwc {
  uint32_t x, c, a;
  __device__ __forceinline__ float operator()() {
    uint64_t t = (uint64_t)a * x + c;
    x = (uint32_t)t;
    c = (uint32_t)(t >> 32);
    asm("" : "+v"(x), "+v"(c));  // two 32-bit values from here on, not a register pair
    return 2.f - __uint_as_float((x >> 9) | 0x3f800000u);
  }
};

// gamma(k) variate (gpu/pro.cu:59-85): Weibull rejection for k<1, Cheng otherwise.  k = 0
// (a particle bunch without longitudinal profile) keeps the reference's powf values
// pow(0, 0) = 1 and pow(z, inf) = 0, inf, 1 for z <, >, = 1, where exp2(y log2 x) would give
// NaN: the loop then draws until z < 1 <= z + e and returns 0, as the reference does.
__device__ float gamma_rnd(float k, Mwc& s) {
  float x;
  if (k < 1) {
    float c = frcp(k);
    float d = (1 - k) * (k > 0 ? fpow(k, frcp(c - 1)) : 1.f);
    float z, e;
    do {
      z = -flog(s());
      e = -flog(s());
      x = k > 0 ? fpow(z, c) : z < 1 ? 0.f : z > 1 ? __builtin_huge_valf() : 1.f;
    } while (z + e < d + x);
  } else {
    float b = k - 1.38629436112f;  // log 4
    float l = fsqrt(2 * k - 1);
    float c = 1 + 1.50407739678f;  // 1 + log 4.5
    float u, v, y, z, r;
    do {
      u = s();
      v = s();
      y = -flog(frcp(v) - 1) * frcp(l);
      x = k * fexp(y);
      z = u * v * v;
      r = b + (k + l) * y - x;
    } while (r < 4.5f * z - c && r < flog(z));
  }
  return x;
}

// Box-Muller: radius from the first draw, angle from the second (gpu/pro.cu:87-95)
__device__ __forceinline__ float gauss_rnd(Mwc& s) {
  float r = fsqrt(-2 * flog(s()));
  return r * __builtin_amdgcn_sinf(s());
}
__device__ __forceinline__ float2 gauss2_rnd(Mwc& s) {
  float r = fsqrt(-2 * flog(s()));
  float u = s();
  return make_float2(r * __builtin_amdgcn_cosf(u), r * __builtin_amdgcn_sinf(u));
}

__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
  float r = frsq(x * x + y * y + z * z);
  x *= r, y *= r, z *= r;
}

// Rotate unit n by the polar angle with 1-cos = cs, sin = si, around a uniform
// azimuth (gpu/pro.cu:111-153); the largest component is swapped to x first, and
// swapped back at the end.  The swaps are selects, so that lanes with different
// largest components do not diverge (ties resolve as in the reference: y only if
// it exceeds both others, z only if it exceeds x and is not below y).
// NORM = false leaves out the final renormalisation.  The turn is perpendicular to
// n and of length si, so |n'|^2 = (1 - cs)^2 |n|^2 + si^2: an error e of |n| becomes
// (1 - cs)^2 e plus rounding, which does not grow over a photon's scatterings; the
// renormalisation (one rsqrt and six operations) is kept for emission, where the
// reference's direction enters once.
template <bool NORM = true>
__device__ __forceinline__ void rotate(float cs, float si, float4& n, Mwc& s) {
  const float x = n.x, y = n.y, z = n.z;
  const float q0 = x * x, q1 = y * y, q2 = z * z;
  const bool sy = (q1 > q2) & (q1 > q0);   // x and y trade places
  const bool sz = !(q1 > q2) & (q2 > q0);  // x and z trade places
  const float nx = sy ? y : sz ? z : x, ny = sy ? x : y, nz = sz ? x : z;
  const float r0 = sy ? q1 : sz ? q2 : q0, r1 = sy ? q0 : q1, r2 = sz ? q0 : q2;
  // the perpendicular pair p1 = a (-ny, nx, 0), p2 = b (-nz, 0, nx) (unit: a, b normalise),
  // turned by 45 degrees into q = (p1 - p2) / |p1 - p2| and p = (p1 + p2) / |p1 + p2|; with
  // p1.p2 = ny nz a b = cc / 2 the norms are sqrt(2 - cc) and sqrt(2 + cc).  The new direction
  // n (1 - cs) + si (cos u q + sin u p) is summed component by component from the factors
  // c1 = si cos u / |p1 - p2| and c2 = si sin u / |p1 + p2|: with E = a (c1 + c2) and
  // F = b (c1 - c2) the turn adds (-ny E + nz F, nx E, -nx F)
  const float a = frsq(r0 + r1), b = frsq(r0 + r2);
  const float cc = 2 * ny * nz * a * b;
  const float u = s();
  const float c1 = si * __builtin_amdgcn_cosf(u) * frsq(2 - cc);
  const float c2 = si * __builtin_amdgcn_sinf(u) * frsq(2 + cc);
  const float E = a * (c1 + c2), F = b * (c1 - c2), omc = 1 - cs;
  float mx = nx * omc + (nz * F - ny * E);
  float my = ny * omc + nx * E;
  float mz = nz * omc - nx * F;
  if (NORM) normalize3(mx, my, mz);
  n.x = sy ? my : sz ? mz : mx;
  n.y = sy ? mx : my;
  n.z = sz ? mx : mz;
}

// ---------------------------------------------------------------- tables

struct Tabs {
  const float4* lay;   // s0, s1, p, t
  const float4* az;    // k1, k2, ra, rb
  const float4* azr;   // 1/k1, 1/k2, 1/(k1 k2), k1 k2
  const float* lr;     // [kLMAX] 1-d tilt reference distances, then [kLMAX] 1/(lr[j] - lr[j-1])
  const int* mcol;
  const uint16_t* is;
  const uint16_t* ls;
  const float4* sc;    // 2 per line: (x, y, r, h), (d, dl, dh, n | max<<16)
  const float2* lp;
  int lp_stride;
};

// wavelength factors of the photon being propagated
struct Wave {
  float la, lk, abl;
  int lay;  // first row of the photon's ice variant in the layer table
  __device__ __forceinline__ float sca(const Tabs& T, int J) const {
    float4 l = T.lay[lay + J];
    return l.x + l.y * la;
  }
  __device__ __forceinline__ float abs(const Tabs& T, int J) const {
    float4 l = T.lay[lay + J];
    return l.z * lk + l.w * abl;
  }
};

// depth shift of the ice layers (gpu/pro.cu:164-264)
template <int TILT>
__device__ __forceinline__ float zshift(const KParams& P, const Tabs& T, const float4& r, float& dz) {
  if constexpr (TILT == 1) {
    // the scalars of this branch read together, so that their loads share one wait
    const float lnx = P.lnx, lny = P.lny, r0 = P.r0, lmin = P.lmin, lrdz = P.lrdz;
    const int lnum = P.lnum, lpts = P.lpts;
    const float lr1 = P.lr[1], lr2 = P.lr[2], lr3 = P.lr[3], lr4 = P.lr[4];
    const float nr = lnx * r.x + lny * r.y - r0;
    // first j >= 1 with nr < lr[j], else lnum - 1: uniform over the reference points, so
    // lanes do not diverge.  Candidates above 4 take the loop, then j = 4..1 are tested
    // from the values read above (IceCube's tilt.par has 6 reference strings: candidates
    // 1..4), in the order of one loop from lnum - 2 down.
    int j = lnum - 1;
    for (int q = lnum - 2; q > 4; q--) j = nr < P.lr[q] ? q : j;
    if (4 <= lnum - 2) j = nr < lr4 ? 4 : j;
    if (3 <= lnum - 2) j = nr < lr3 ? 3 : j;
    if (2 <= lnum - 2) j = nr < lr2 ? 2 : j;
    if (1 <= lnum - 2) j = nr < lr1 ? 1 : j;
    const float z = (r.z - lmin) * lrdz;
    const int i = j - 1, k = min(max(rd(z), 0), lpts);
    // from global memory (L1/L2-resident; measured faster than an LDS copy here, which
    // would cost occupancy), with 32-bit offsets
    const uint32_t e = i * T.lp_stride + k;
    const float2 pi = ldg(P.lp, e), pj = ldg(P.lp, e + T.lp_stride);
    const float rdw = T.lr[kLMAX + j];  // 1 / (lr[j] - lr[j-1]), prepared on the host
    float wx = (nr - T.lr[i]) * rdw, wy = (T.lr[j] - nr) * rdw;
    float qx = pj.x * wx + pi.x * wy;
    float qy = pj.y * wx + pi.y * wy;
    if (P.vthk != 0) {
      asm volatile("");  // a branch on the uniform flag, not a select that always divides
      dz = fdiv(dz, 1 - lrdz * qy);
    }
    return qx + qy * (z - k);
  } else if constexpr (TILT == 2) {
    float qx = P.cb00 * r.x + P.cb10 * r.y;
    float qy = P.cb01 * r.x + P.cb11 * r.y;
    qx = (qx - P.mmin0) / P.mstp0;
    qy = (qy - P.mmin1) / P.mstp1;
    {  // extrapolate outside the hexagon of the map
      float x = qx, y = qy, xy;
      int n0 = P.mnum0 - 1, n1 = P.mnum1 - 1, c0 = P.mcut0 - 1, c1 = P.mcut1 - 1;
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
    }
    int nx = min(max(rd(qx), 0), P.mnum0);
    int ny = min(max(rd(qy), 0), P.mnum1);
    float fx = qx - nx, fy = qy - ny;
    float qrx, qry;
    int qnx, qny;
    if (fx > fy)
      qrx = fx, qry = fy, qnx = nx + 1, qny = ny;
    else
      qrx = fy, qry = fx, qnx = nx, qny = ny + 1;
    auto col = [&](int a, int b) {
      return a >= 0 && a < P.mnum0 && b >= 0 && b < P.mnum1 ? T.mcol[a * kCTY + b] : -1;
    };
    int cx = col(nx, ny), cy = col(qnx, qny), cz = col(nx + 1, ny + 1);
    float wx = 1 - qrx, wy = qrx - qry, wz = qry;
    float z = (r.z - P.lmin) * P.lrdz;
    int k = min(max(rd(z), 0), P.lpts);
    float ax = 0, ay = 0;
    if (cx >= 0) {
      float2 v = T.lp[cx * T.lp_stride + k];
      ax += v.x * wx, ay += v.y * wx;
    }
    if (cy >= 0) {
      float2 v = T.lp[cy * T.lp_stride + k];
      ax += v.x * wy, ay += v.y * wy;
    }
    if (cz >= 0) {
      float2 v = T.lp[cz * T.lp_stride + k];
      ax += v.x * wz, ay += v.y * wz;
    }
    if (P.vthk != 0) {
      asm volatile("");  // a branch on the uniform flag, not a select that always divides
      dz = fdiv(dz, 1 - P.lrdz * ay);
    }
    return ax + ay * (z - k);
  } else {
    return 0.f;
  }
}

// Lengths of a straight path of length L from layer coordinate z (layer I)
// along a direction with |cos zenith| = anz, split over the layers it crosses:
// the partial first layer, whole layers in between, the rest in the last
// layer -- the split of gpu/pro.cu:681-696.  f(k, l) is called per layer.
template <class F>
__device__ __forceinline__ void split_layers(const KParams& P, float z, int I, int sign, float anz, float edh, float L,
                                             F f) {
  const float y = z + sign * anz * L / edh;
  const int J = min(max(rn(y), 0), P.size - 1);
  if (I == J) {
    f(I, L);
    return;
  }
  const float full = edh / anz;
  float l = (0.5f - sign * (z - I)) * full, used = 0;
  for (int k = I; k != J; k += sign) {
    if (l > L - used) l = L - used;
    f(k, l);
    used += l;
    l = full;
  }
  f(J, fmaxf(L - used, 0.f));
}

// ---------------------------------------------------------------- lane state

struct Photon {
  float4 r, n;  // n.w = 1/group velocity
  float TOT, SCA;
  int old, ofla;
  uint32_t nhw;     // hits of this photon so far (low 16 bits), wavelength bin << 16
  float clear;      // lower bound of the horizontal distance to every string cylinder (0: unknown)
  Wave w;
};

// Emit photon i: pick its bunch, a wavelength bin, and the source-specific
// start and direction (gpu/pro.cu:334-477).  Returns the starting DOM (fla).
__device__ int emit(const KParams& P, uint32_t i, Mwc& s, Photon& ph, uint32_t& seg) {
  // bunch: last b with pstart[b] <= i
  uint32_t lo = 0, hi = P.nb;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (P.pstart[mid] <= i) lo = mid; else hi = mid;
  }
  const ppc_hip_bunch& p = P.pz[lo];

  int j = min(rd(kWNUM * s()), kWNUM - 1);
  const ppc_hip_wv wv = P.wv[j];
  float4 r = make_float4(p.r[0], p.r[1], p.r[2], p.r[3]);
  float4 n = make_float4(p.n[0], p.n[1], p.n[2], wv.ocm);
  float l = p.n[3];
  int fla, ofla;

  if (p.type > 0) {  // flashers and standard candles
    fla = p.u.l.fla, ofla = p.u.l.ofla;
    const float ka = p.u.l.ka, up = p.u.l.up, fldr = p.u.l.fldr;
    if (l == 0) {  // LED ring: uniform azimuth or N-fold pattern
      float xi = s();
      if (fldr < 0)
        xi *= 2 * FPI;
      else {
        int rr = rd(fldr / 360) + 1;
        int ss = rd(xi * rr);
        xi = (fldr + ss * 360 / rr) * (FPI / 180.f);
      }
      sincosf(xi, &r.y, &r.x);
      float cu;
      sincosf(up, &n.z, &cu);
      n.x = cu * r.x;
      n.y = cu * r.y;
    }
    if (ka != 0) {
      float xi;
      if (ka > 999.f) {  // lab-measured beam profile
        xi = s();
        float s1 = 0.0191329f, s2 = 0.0686944f, C1 = 0.583583f, C2 = 0.967762f;
        if (p.type == 1) s1 = 0.0196242f, s2 = 0.0750278f, C1 = 0.606833f, C2 = 0.960863f;
        if (p.type == 2) s1 = 0.0185603f, s2 = 0.0624481f, C1 = 0.553192f, C2 = 0.974129f;
        xi = xi < C1 ? s1 * fabsf(gauss_rnd(s)) : xi < C2 ? -s2 * flog(s()) : 2.f;
        if (xi >= 2.f) xi = 2 * (1 - fsqrt(s()));
      } else if (ka > 0)  // 2-d gaussian
        xi = fminf(-ka * flog(1 - (1 - s()) * (1 - expf(-2 / ka))), 2.f);
      else if (ka == -1)  // isotropic
        xi = 2 * (1 - s());
      else  // linear cone
        xi = 2 * (1 - fsqrt(s()));
      rotate(xi, fsqrt(xi * (2 - xi)), n, s);
    }
    if (ka > 999.f) {  // start on the flasher board inside the DOM
      const float FLZ = 0.07735f, FLR = 0.119f - 0.008f, FLB = 0.03396f;
      if (l == 0) r.x *= FLR, r.y *= FLR, r.z = FLZ, r.w = 0;
      float b = r.x * n.x + r.y * n.y + r.z * n.z;
      float c = FLZ * FLZ + FLR * FLR - OMR * OMR;
      float t = sqrtf(b * b - c) - b;
      r.x += t * n.x, r.y += t * n.y, r.z += t * n.z, r.w += t * P.ocv;
      if (fabsf(r.z) < FLB)
        ofla = -2;
      else if (l == 0 && r.z < 0 && s() < 0.9f)
        ofla = -3;
      if (ka > 1050.f) {  // cable shadow
        const float rc = 0.023f, dc = OMR + rc;
        float phc = (ka - 1080.f) * (FPI / 180.f);
        float drx = r.x - dc * cosf(phc), dry = r.y - dc * sinf(phc);
        float a = n.x * n.x + n.y * n.y;
        if (a > 0) {
          float b2 = n.x * drx + n.y * dry;
          float c2 = drx * drx + dry * dry - rc * rc;
          float D = b2 * b2 - a * c2;
          if (D >= 0 && (-b2 + sqrtf(D)) / a > 0) ofla = -4;
        }
      }
      if (l == 0)
        r.x += p.r[0], r.y += p.r[1], r.z += p.r[2], r.w += p.r[3];
      else {
        float4 d = P.dom[fla];
        r.x += d.x, r.y += d.y, r.z += d.z, r.w += p.r[3];
      }
    } else if (fla >= 0) {  // start on the DOM surface
      r.x = OMR * n.x + p.r[0], r.y = OMR * n.y + p.r[1], r.z = OMR * n.z + p.r[2];
      r.w = OMR * P.ocv + p.r[3];
    } else
      r = make_float4(p.r[0], p.r[1], p.r[2], p.r[3]);
  } else {  // particles: muon segments and cascades
    fla = -1, ofla = -1;
    const float pa = p.u.p.a, pb = p.u.p.b, beta0 = p.u.p.beta, tau = p.u.p.tau;
    if (l > 0)
      l *= s();
    else
      l = pb * gamma_rnd(pa, s);
    if (l > 0) {
      r.x += l * n.x, r.y += l * n.y, r.z += l * n.z;
      r.w += l * P.ocv / beta0;
    }
    if (tau > 0) {  // isotropic delayed emission
      r.w -= tau * flog(s());
      float xi = 1 - s();
      rotate(xi, fsqrt(xi * (2 - xi)), n, s);
    } else {
      float cs = wv.coschr, si = wv.sinchr;
      if (p.f < s()) {  // cascade particle directions
        const float a = 0.39f, b = 2.61f;
        const float I = 1 - expf(-b * exp2f(a));
        float xi = fminf(fpow(-flog(1 - s() * I) / b, 1 / a), 2.f);
        rotate(xi, fsqrt(xi * (2 - xi)), n, s);
      } else {
        float beta = beta0;
        if (p.type < 0) {  // muon end point
          const float ar = 0.26f * 0.9216f / 0.105658389f;
          float dx = ar * (p.n[3] - l);
          beta = sqrtf(dx * (2 + dx)) / (1 + dx);
          if (beta >= cs) r.w -= P.ocv * (sqrtf(dx * dx + 1) - asinhf(1 / dx) - dx) / ar;
        }
        if (beta < 1) {
          float xi = cs / beta;
          if (xi <= 1) {
            float sx = sqrtf(1 - xi * xi);
            if (p.type < 0 && sx < s() * si) ofla = -3;
            cs = xi, si = sx;
          } else
            ofla = -2;
        }
      }
      rotate(1 - cs, si, n, s);  // Cherenkov cone
    }
  }
  ph.r = r, ph.n = n;
  ph.nhw = (uint32_t)wv.wvl << 16;
  seg = p.q;
  ph.w.la = wv.la, ph.w.lk = wv.lk, ph.w.abl = wv.abl;
  ph.w.lay = P.bvar ? (int)P.bvar[lo] * P.size : 0;
  ph.ofla = ofla;
  return fla;
}

// ---------------------------------------------------------------- kernel

// The parameter block is read through a pointer into device memory: uniform
// loads become scalar loads the compiler can repeat, where a by-value argument
// of this size would be held in SGPRs and spilled into VGPR lanes.
// Occupancy: the kernel is VALU-issue bound with divergent branches, so more
// waves per SIMD keep the VALU busier; at 7 waves the register budget is 72 VGPRs
// (a few spills: reloaded at emission, after a DOM search and on the hit paths, not
// on the plain step), measured faster than 4, 5, 6 (no spills) and 8.
// ADJ: replay the photons of adjoint records (each from its recorded MWC state)
// and accumulate d(LLH)/d(per-layer absorption, scattering) instead of hits:
// along every path segment the score of the path density, -nr*d(abs)/da*l for
// absorption and d(sca)/db*(1/sca at a scattering - l) for scattering, weighted
// by the LLH weight of the hits the photon still makes after the segment.
template <int TILT, bool EXT, bool COUNT, bool ADJ>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(kWavesPerSimd))) void propagate(
    const KParams* __restrict__ Pp) {
  const KParams& P = *Pp;
  // start of the in-kernel time span, once per workgroup (its waves start together)
  if (threadIdx.x == 0) atomicMax(&P.ctr->t_start_c, ~wall_clock64());
  extern __shared__ uint4 lds[];
  for (int k = threadIdx.x; k < P.image_words; k += blockDim.x) lds[k] = P.image[k];
  // per lane, after the image: the photon's index, its segment and the MWC state it started
  // from (written at emission, read back only for hits, records and the adjoint's second pass)
#define info (reinterpret_cast<uint32_t*>(lds + P.image_words) + threadIdx.x)
  // (four words per lane, a stride of blockDim.x apart: photon, segment, x, c)
  float* gacc = reinterpret_cast<float*>(lds + P.image_words + blockDim.x);  // ADJ: [2][size] block sums
  if (ADJ)
    for (int k = threadIdx.x; k < 2 * P.size; k += blockDim.x) gacc[k] = 0.f;
  __syncthreads();

  const char* base = reinterpret_cast<const char*>(lds);
  Tabs T;
  T.lay = reinterpret_cast<const float4*>(base + P.off_lay);
  T.az = reinterpret_cast<const float4*>(base + P.off_az);
  T.azr = T.az + P.size;
  T.lr = reinterpret_cast<const float*>(base + P.off_lr);
  T.mcol = reinterpret_cast<const int*>(base + P.off_mcol);
  T.is = reinterpret_cast<const uint16_t*>(base + P.off_is);
  T.ls = reinterpret_cast<const uint16_t*>(base + P.off_ls);
  T.sc = reinterpret_cast<const float4*>(base + P.off_sc);
  T.lp = P.lp_in_lds ? reinterpret_cast<const float2*>(base + P.off_lp) : P.lp;

  T.lp_stride = P.lp_stride;

  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = __lane_id();
  Mwc s;
  {
    uint2 st = P.rs[gid];
    s.x = st.x, s.c = st.y, s.a = P.rm[gid];
  }

  Photon ph;
  ph.TOT = 0, ph.SCA = 0, ph.old = -1, ph.ofla = -1;
  ph.r = make_float4(0.f, 0.f, 0.f, 0.f), ph.n = make_float4(0.f, 0.f, 1.f, 0.f);
  ph.clear = 0.f, ph.nhw = 0, ph.w.la = 0.f, ph.w.lk = 0.f, ph.w.abl = 0.f, ph.w.lay = 0;
  bool live = false;
  bool done = false;  // no more photons to claim (wave-uniform)
  uint32_t next_static = gid;
  // photons this wave has claimed but not started: [res_next, res_end) (wave-uniform);
  // one atomic claims a chunk, so the global counter sees one request per chunk
  uint32_t res_next = 0, res_end = 0;
  const uint32_t chunk = P.chunk;
  // steps, photons and killed photons of the wave (uniform: scalar registers), event counts
  // of the lane (COUNT builds only); non-finite photons (rare) go straight to the launch counter
  uint32_t n_steps = 0, n_photons = 0, n_killed = 0;
  uint32_t n_cross = 0, n_str = 0, n_dom = 0, n_scat = 0;
  // ADJ: each record's photon runs twice from its start state: pass 0 sums the LLH
  // weights of its hits (Gsum), pass 1 scores the path with G, the weight of the hits
  // still to come
  float G = 0.f, Gsum = 0.f;
  uint32_t nw = 0;  // weighted hits of the photon: counted in pass 0, counted down in pass 1
  bool pass1 = false, pend = false;
  const uint32_t nwork = ADJ ? P.adj_nrec : P.num;

  while (true) {
    // -------- refill idle lanes --------
    if (ADJ && done && __any(pend)) done = false;  // a summed photon still has its scoring pass to run
    if (!done) {
      const uint64_t idle = __ballot(!live);
      const int nidle = __popcll(idle);
      if (nidle >= P.refill || nidle == __popcll(__ballot(true))) {
        uint32_t i = 0xffffffffu;
        if (ADJ) {  // one record after the other per lane
          if (!live && pend) {  // the scoring pass of the photon just summed
            s.x = info[2 * blockDim.x], s.c = info[3 * blockDim.x];
            i = info[0];
            G = Gsum, pass1 = true, pend = false;  // nw carries over from pass 0
          } else if (!live && next_static < nwork) {
            // (an exhausted lane's counter stays put: advancing it at every refill would
            // wrap the 32-bit index around to work already done)
            const ppc_hip_adj_record rc = P.adj_rec[next_static];
            next_static += P.lanes;
            s.x = rc.x0, s.c = rc.c0, s.a = P.rm[rc.lane];
            i = rc.photon;
            G = 0.f, Gsum = 0.f, nw = 0, pass1 = false;
          }
          done = !__any(next_static < nwork || pend);
        } else if (P.deterministic) {
          // photon next_static, then next_static + lanes, ...; a lane past the end keeps its
          // counter (advancing it at every refill would wrap it around to photons already run)
          if (!live && next_static < P.num) i = next_static, next_static += P.lanes;
          done = !__any(next_static < P.num);
        } else {
          const uint32_t rank = lanes_below(idle);
          const uint32_t avail = res_end - res_next;
          if (avail >= (uint32_t)nidle) {
            if (!live) i = res_next + rank;
            res_next += nidle;
          } else {  // take the rest of the reservoir, then a new chunk from the global counter
            uint32_t first = 0;
            const int leader = __ffsll((unsigned long long)idle) - 1;
            if ((int)lane == leader) first = atomicAdd(&P.ctr->next, chunk);
            first = __builtin_amdgcn_readfirstlane(__shfl(first, leader));
            if (!live) i = rank < avail ? res_next + rank : first + (rank - avail);
            const uint32_t used = nidle - avail;
            if (first >= P.num) {  // the launch is exhausted: the reservoir was the last work
              res_next = res_end = 0;
              done = true;
            } else {
              res_next = first + used;
              res_end = (unsigned long long)first + chunk > P.num ? P.num : first + chunk;
              if (res_next > res_end) res_next = res_end;
            }
          }
        }
        const bool start = !live && i < P.num;
        n_photons += __popcll(__ballot(start));
        if (start) {
          if (!ADJ && P.photon_seed != 0) {
            // a start state of the photon's own, a hash of (seed, i) on its lane's multiplier:
            // launches with the same seed draw the same numbers for photon i whatever the
            // other photons did (common random numbers); 0 < c < a, x odd
            const uint64_t h = mix64(P.photon_seed + (uint64_t)(i + 1u) * 0x9E3779B97F4A7C15ull);
            s.x = (uint32_t)h | 1u;
            s.c = 1u + __umulhi((uint32_t)(h >> 32), s.a - 2u);
          }
          const uint32_t x0 = s.x, c0 = s.c;
          if (!ADJ && P.rec_stride != 0 && i % P.rec_stride == 0 && i / P.rec_stride < P.rec_max) {
            ppc_hip_record rc;
            rc.photon = i, rc.x = s.x, rc.c = s.c, rc.a = s.a;
            P.rec[i / P.rec_stride] = rc;
          }
          ph.clear = 0.f;
          uint32_t seg;
          int fla = emit(P, i, s, ph, seg);
          info[0] = i, info[blockDim.x] = seg, info[2 * blockDim.x] = x0, info[3 * blockDim.x] = c0;
          ph.TOT = -flog(s());
          ph.SCA = 0;
          ph.old = fla;
          if (ph.ofla < -1) {  // killed at emission: the reference still draws SCA
            s();
            ph.TOT = 0;
          } else
            live = true;
        }
        n_killed += __popcll(__ballot(start && !live));
      }
    }
    if (done && !__any(live)) break;
    n_steps += __popcll(__ballot(live));
    // Every lane runs the step, a lane without a photon on a state that changes nothing:
    // it draws nothing (SCA stays 0), searches and records nothing, its zero optical
    // depths walk no layer and end in the death test (a 0/0 there too).  No per-lane
    // region around the step and one loop back-edge: with either, the register allocator
    // kept the photon state in two register sets and copied it between them every step.
    {
    const bool alive = live;
    // -------- one transport step (gpu/pro.cu:486-889) --------
    float4 r = ph.r, n = ph.n;
    float TOT = ph.TOT, SCA = ph.SCA;
    int old = ph.old;
    const int ofla = ph.ofla;
    const Wave w = ph.w;

    if (alive && SCA == 0) SCA = -flog(s());  // old keeps the DOM of the previous step / the emitter
    float sca, tot;
    const int sign = n.z < 0 ? -1 : 1;
    const float anz = sign * n.z;
    float edh = P.dh;
    float z = r.z - zshift<TILT>(P, T, r, edh);
    float nr = 1.f;
    int I, J;
    {
      z = (z - P.hmin) * P.rdh;
      I = min(max(rn(z), 0), P.size - 1);
      if (TOT > 0) {  // absorption anisotropy
        float n1 = P.azx * n.x + P.azy * n.y;
        float n2 = -P.azy * n.x + P.azx * n.y;
        float n3 = n.z;
        if (EXT && P.fr != 0) {
          float4 a = T.az[I];
          float k1 = a.x, k2 = a.y;
          float kz = frcp(k1 * k2);
          float s1 = n1 * n1, l1 = k1 * k1;
          float s2 = n2 * n2, l2 = k2 * k2;
          float s3 = n3 * n3, l3 = kz * kz;
          float B2 = nr / l1 + nr / l2 + nr / l3;
          float nB = s1 / l1 + s2 / l2 + s3 / l3;
          float An = s1 * l1 + s2 * l2 + s3 * l3;
          nr = fpow((B2 - nB) * An / 2, P.fr);
        }
        n1 *= P.k1, n2 *= P.k2, n3 *= P.kz;
        const float q = n1 * n1 + n2 * n2 + n3 * n3, rq = frsq(q);
        if (EXT && P.fr != 0) {
          nr *= rq;
          TOT = fdiv(TOT, nr);
        } else {
          nr = rq;
          TOT *= q * rq;  // 1/nr = sqrt(q) = q rsq(q)
        }
      }
      float ahx = edh * (0.5f - sign * (z - I));
      J = I;
      sca = 0, tot = 0;
      float ais = SCA, aia = TOT;
      const float ranz = frcp(anz);
      const int jend = sign < 0 ? 0 : P.size - 1;  // the last layer in the direction of travel
      float bend;  // absorption coefficient of the layer the walk ends in
      while (true) {  // layer walk
        const float bs = w.sca(T, J), ba = w.abs(T, J);
        float dsca = fdiv(ais, bs), dtot = fdiv(aia, ba);
        float dist = fminf(dsca, dtot) * anz;
        if ((J != jend) & (dist > ahx)) {
          float del = ahx * ranz;
          sca += del, tot += del;
          ais -= bs * del, aia -= ba * del;
          ahx = edh;
          J += sign;
          if (COUNT) n_cross++;
        } else {
          sca += dsca, tot += dtot;
          bend = ba;
          break;
        }
      }
      if (sca < 0) sca = 0;
      if (tot < 0) tot = 0;
      if (tot < sca)
        sca = tot, tot = 0;
      else
        tot = (tot - sca) * bend;
    }

    int om = -1;
    float del = sca, sqd = 0;
    float hi = sca, hf = 0;
    const float np = 1 - n.z * n.z;
    // DOM and hole-ice search in the cells overlapping the step, unless the step cannot
    // reach any string cylinder: its horizontal length is below the clearance of its
    // start, a lower bound of the distance to every cylinder that the photon carries from
    // step to step (less the horizontal length of each step) and renews from the clearance
    // map when it runs short.  The cylinder test of every string would fail on such a
    // step, so the skip changes no result.
    bool search = true;
    if (P.clear_map) {
      const float reach = sca * fsqrt(np);
      if (alive & (ph.clear <= reach)) {
        ph.clear = 0.f;
        // the map's parameters read together (one wait for their scalar loads), and the
        // bounds tested without short-circuit branches
        const float* cmap = P.clear_map;
        const float x0 = P.cm_x0, y0 = P.cm_y0, rs = P.cm_rstep;
        const int nx = P.cm_nx, ny = P.cm_ny;
        const float gx = (r.x - x0) * rs, gy = (r.y - y0) * rs;
        const bool in = (gx >= 0.f) & (gy >= 0.f) & (gx < (float)nx) & (gy < (float)ny);
        if (in) ph.clear = ldg(cmap, (uint32_t)((int)gy * nx + (int)gx));
      }
      if (ph.clear > reach) {
        search = false;
        // a curved path (birefringence) ends off the straight line: renew every step
        ph.clear = (EXT && P.sum > 0) ? 0.f : ph.clear - reach;
      } else
        ph.clear = 0.f;
    }
    search &= alive;
    if (search) {
      float rfx = r.x + sca * n.x, rfy = r.y + sca * n.y;
      float pix = P.cb00 * r.x + P.cb10 * r.y, piy = P.cb01 * r.x + P.cb11 * r.y;
      float pfx = P.cb00 * rfx + P.cb10 * rfy, pfy = P.cb01 * rfx + P.cb11 * rfy;
      float lox = fminf(pix, pfx) - P.rx, hix = fmaxf(pix, pfx) + P.rx;
      float loy = fminf(piy, pfy) - P.rx, hiy = fmaxf(piy, pfy) + P.rx;
      int xlx = min(max(rn((lox - P.cl0) * P.crst0), 0), P.cn0);
      int xhx = max(min(rn((hix - P.cl0) * P.crst0), P.cn0 - 1), -1);
      int xly = min(max(rn((loy - P.cl1) * P.crst1), 0), P.cn1);
      int xhy = max(min(rn((hiy - P.cl1) * P.crst1), P.cn1 - 1), -1);
      for (int ci = xlx; ci <= xhx; ci++)
        for (int cj = xly; cj <= xhy; cj++)
          for (uint32_t k = T.is[ci * kCY + cj]; k != 0x8000u;) {
            const uint32_t m = T.ls[k];
            k = (m & 0x8000u) ? 0x8000u : k + 1;
            const float4 la = T.sc[2 * (m & 0x7fffu)];
            if (COUNT) n_str++;
            float b = 0, c = 0, dr;
            dr = la.x - r.x;
            b += n.x * dr, c += dr * dr;
            dr = la.y - r.y;
            b += n.y * dr, c += dr * dr;
            float D = b * b - (c - la.z * la.z) * np;
            if (D >= 0) {
              D = fsqrt(D);
              float h1 = b - D, h2 = b + D;
              if (h2 >= 0 && h1 <= sca * np) {
                const float4 lb = T.sc[2 * (m & 0x7fffu) + 1];
                if (np > XXX) {
                  const float rnp = frcp(np);
                  h1 *= rnp, h2 *= rnp;
                  if (h1 < 0) h1 = 0;
                  if (h2 > sca) h2 = sca;
                } else
                  h1 = 0, h2 = sca;
                h1 = r.z + n.z * h1, h2 = r.z + n.z * h2;
                float zl = n.z > 0 ? h1 : h2, zh = n.z > 0 ? h2 : h1;
                const uint32_t bits = __float_as_uint(lb.w);
                const int sn = bits & 0xffff, smax = bits >> 16;
                int n1 = sn + min(smax + 1, max(0, ru(-(zh - lb.y - la.w) * lb.x)));
                int n2 = sn + max(-1, min(smax, rd(-(zl - lb.z - la.w) * lb.x)));
                for (int l = n1; l <= n2; l++) {
                  if (l == old) continue;
                  if (COUNT) n_dom++;
                  const float4 dm = ldg(P.dom, (uint32_t)l);
                  const float F = P.uniform_dom ? 1.f : ldg(P.domF, (uint32_t)l);
                  float a = 0, bb = 0, cc = 0, d2;
                  const float f = F > 0 ? sq(frcp(F)) : 0;
                  d2 = dm.x - r.x;
                  a += n.x * n.x, bb += n.x * d2, cc += d2 * d2;
                  d2 = dm.y - r.y;
                  a += n.y * n.y, bb += n.y * d2, cc += d2 * d2;
                  d2 = dm.z - r.z;
                  a += f * n.z * n.z, bb += f * n.z * d2, cc += f * d2 * d2;
                  if (a > 0) {
                    const float ra = frcp(a);
                    bb *= ra;
                    const float xR = (l == ofla ? 1.f : P.xR);
                    float DD = bb * bb - (cc - sq(xR * dm.w)) * ra;
                    if (DD >= 0) {
                      sqd = fsqrt(DD);
                      float h = bb - sqd;
                      if (F > 0 || fabsf(d2 - n.z * h) < -xR * F * dm.w) {
                        sqd *= frcp(xR);
                        h = bb - sqd;
                        if (h > 0 && h <= del) om = l, del = h;
                      }
                    }
                  }
                }
              }
            }
            if (EXT && P.hr > 0 &&
                (P.hifl == 0 || (P.hifl < 0 && r.w < -P.hifl) || (P.hifl > 0 && r.w > P.hifl))) {
              float D2 = b * b - (c - P.hr2) * np;
              if (D2 > 0) {
                D2 = fsqrt(D2);
                float h1 = b - D2, h2 = b + D2;
                if (h2 >= 0 && h1 <= sca * np) {
                  if (np > XXX) {
                    h1 /= np, h2 /= np;
                    if (h1 < 0) h1 = 0;
                    if (h2 > sca) h2 = sca;
                  } else
                    h1 = 0, h2 = sca;
                  if (h1 < hi && h2 > 0.0031622776f * P.hr) hi = h1, hf = h2;
                }
              }
            }
          }
    }

    const float fin = fminf(del, hi);
    const bool hole = fin < sca;
    if (hole) {  // step cut short by a DOM or the hole-ice column: spend the optical depths
      float xs = 0, xa = 0;
      const float y = z + n.z * fin / edh;
      J = min(max(rn(y), 0), P.size - 1);
      if (I == J)
        xs = fin * w.sca(T, I), xa = fin * w.abs(T, I);
      else {
        const float h = 0.5f - sign * (z - I);
        const float g = 0.5f + sign * (y - J);
        xs = h * w.sca(T, I) + g * w.sca(T, J);
        xa = h * w.abs(T, I) + g * w.abs(T, J);
        for (int i = I + sign; i != J; i += sign) xs += w.sca(T, i), xa += w.abs(T, i);
        const float f = edh / anz;
        xs *= f, xa *= f;
      }
      if (xs < 0) xs = 0;
      if (xa < 0) xa = 0;
      SCA -= xs, TOT -= xa;
      TOT *= nr;
      if (EXT && hi < del) {  // inside the hole ice
        float fin2 = fminf(hi + fminf(SCA / P.hs, TOT / P.ha), hf);
        if (om != -1) {
          if (fin2 < del) del = fin2, om = -1;
        } else {
          if (fin2 < del)
            del = fin2;
          else
            del = fminf(fin2, sca);
        }
        fin2 = del - hi;
        SCA -= fin2 * P.hs, TOT -= fin2 * P.ha;
      }
    } else
      SCA = 0, TOT = tot * nr;

    if (ADJ && alive && G != 0.f) {  // path-length scores of this segment
      const float ca = -G * nr * w.lk, cb = -G * w.la;
      split_layers(P, z, I, sign, anz, edh, del, [&](int k, float l) {
        atomicAdd(&gacc[k], ca * P.adj_dpda[k] * l);
        atomicAdd(&gacc[P.size + k], cb * P.adj_ds1db[k] * l);
      });
    }

    r.x += del * n.x, r.y += del * n.y, r.z += del * n.z, r.w += del * n.w;

    if (alive && (!isfinite(TOT) || !isfinite(SCA))) {  // rare: counted straight into the launch counter
      TOT = 0, om = -1;
      atomicAdd(&P.ctr->nonfinite, 1ull);
    }

    if (ADJ && om != -1) {
      if (om == ofla)
        TOT = 0;
      else {
        const uint32_t seg = info[blockDim.x];
        if (seg < P.adj_nseg) {
          const float wh = P.adj_weight[(size_t)seg * P.gsize + om];
          if (!pass1)
            Gsum += wh, nw++;
          else if (nw > 0)
            G = --nw == 0 ? 0.f : G - wh;  // exactly 0 after the last weighted hit
        }
        if (P.xR == 1)
          TOT = 0;
        else
          old = om;
      }
    } else if (om != -1) {  // DOM hit
      if (om == ofla)
        TOT = 0;
      else {
        ppc_hip_hit h;
        h.i = om, h.t = r.w, h.n = info[blockDim.x], h.z = ph.nhw >> 16;
        h.pth = acosf(fminf(fmaxf(n.z, -1.f), 1.f));  // |n| is 1 up to rounding
        h.pph = atan2f(n.y, n.x);
        sqd *= P.xR - 1;
        const float4 dm = P.dom[om];
        const float F = P.uniform_dom ? 1.f : P.domF[om];
        float dx = dm.x - r.x + sqd * n.x;
        float dy = dm.y - r.y + sqd * n.y;
        float dz = dm.z - r.z + sqd * n.z;
        h.dth = acosf(fminf(fmaxf(dz / (P.xR * dm.w * F), -1.f), 1.f));
        h.dph = atan2f(dy, dx);
        // wave-level compaction of the lanes that hit in this step
        const uint64_t hm = __ballot(true);
        const int leader = __ffsll((unsigned long long)hm) - 1;
        uint32_t first = 0;
        if ((int)lane == leader) {
          first = atomicAdd(&P.ctr->hidx, (uint32_t)__popcll(hm));
        }
        first = __shfl(first, leader);
        const uint32_t slot = first + lanes_below(hm);
        if (slot < P.hnum) {
          P.hits[slot] = h;
          P.hit_photon[slot] = info[0];
        }
        ph.nhw++;
        if (P.xR == 1)
          TOT = 0;
        else
          old = om;
      }
    }

    // (written so that NaN ends here too: an idle lane's 0/0 in a layer without absorption)
    if (!(TOT >= XXX)) {
      TOT = 0;
      live = false;
      if (!alive) {
      } else if (ADJ) {
        if (!pass1 && nw > 0) pend = true;
        if (pass1 && nw != 0) atomicAdd(&P.ctr->nonfinite, 1ull);  // the scoring pass lost a hit of the summing pass
      } else if (P.hit_photons && (ph.nhw & 0xffffu)) {  // start state of a photon that hit, for the adjoint
        const uint64_t em = __ballot(true);
        const int leader = __ffsll((unsigned long long)em) - 1;
        uint32_t first = 0;
        if ((int)lane == leader) first = atomicAdd(&P.ctr->nphot, (uint32_t)__popcll(em));
        first = __shfl(first, leader);
        const uint32_t slot = first + lanes_below(em);
        if (slot < P.hit_photons_cap) {
          ppc_hip_adj_record rc;
          rc.photon = info[0], rc.x0 = info[2 * blockDim.x], rc.c0 = info[3 * blockDim.x], rc.lane = gid;
          P.hit_photons[slot] = rc;
        }
      }
    } else {
      if (EXT && P.sum > 0 && fin > 0) {  // birefringence (gpu/pro.cu:757-836)
        float ra, rb;
        if (I == J) {
          float4 a = T.az[I];
          ra = fin * a.z, rb = fin * a.w;
        } else {
          const float y = z + n.z * fin / edh;
          const float h = 0.5f - sign * (z - I);
          const float g = 0.5f + sign * (y - J);
          float4 aI = T.az[I], aJ = T.az[J];
          ra = h * aI.z + g * aJ.z;
          rb = h * aI.w + g * aJ.w;
          for (int i = I + sign; i != J; i += sign) {
            float4 a = T.az[i];
            ra += a.z, rb += a.w;
          }
          const float f = edh / anz;
          ra *= f, rb *= f;
        }
        if (ra < 0) ra = 0;
        if (rb < 0) rb = 0;
        const float dot = P.azx * n.x + P.azy * n.y;
        const float sdt = sqrtf(1 - dot * dot);
        const float cdt = fabsf(dot);
        const float* B = P.bfr;
        const float sra = sqrtf(ra);
        float sx = sra * fmaxf(0.f, B[0] * fexp(-B[1] * fpow(atanf(B[3] * sdt), B[2])));
        float sy = sra * fmaxf(0.f, B[4] * fexp(-B[5] * fpow(atanf(B[7] * sdt), B[6])));
        float mx = rb * fmaxf(0.f, B[8] * atanf(B[11] * sdt * cdt) * fexp(-B[9] * sdt + B[10] * cdt));
        float qx = P.azx - dot * n.x, qy = P.azy - dot * n.y, qz = -dot * n.z;
        if (dot < 0) qx = -qx, qy = -qy, qz = -qz;
        float qn = qx * qx + qy * qy + qz * qz;
        if (qn > 0) {
          float f = frsq(qn);
          qx *= f, qy *= f, qz *= f;
        } else
          qx = 0, qy = 0, qz = 1;
        const float px = n.y * qz - n.z * qy, py = n.z * qx - n.x * qz, pz = n.x * qy - n.y * qx;
        float2 g2 = gauss2_rnd(s);
        float dnx = sx * g2.x + mx;
        float dny = sy * g2.y;
        if (del <= hi) {  // curved path within the step
          g2 = gauss2_rnd(s);
          const float r3 = 0.57735026919f;
          float vnx = dnx + sx * g2.x * r3;
          float vny = dny + sy * g2.y * r3;
          vnx /= 2, vny /= 2;
          float den = vnx * vnx + vny * vny;
          float dnz = -den;
          den = 2 / (1 + den);
          vnx *= den, vny *= den, dnz *= den;
          float mxv = n.x + qx * vnx + px * vny + n.x * dnz;
          float myv = n.y + qy * vnx + py * vny + n.y * dnz;
          float mzv = n.z + qz * vnx + pz * vny + n.z * dnz;
          normalize3(mxv, myv, mzv);
          float dl = del - fin / 2;
          r.x += dl * (mxv - n.x), r.y += dl * (myv - n.y), r.z += dl * (mzv - n.z);
        }
        dnx /= 2, dny /= 2;
        float den = dnx * dnx + dny * dny;
        float dnz = -den;
        den = 2 / (1 + den);
        dnx *= den, dny *= den, dnz *= den;
        n.x += qx * dnx + px * dny + n.x * dnz;
        n.y += qy * dnx + py * dny + n.y * dnz;
        n.z += qz * dnx + pz * dny + n.z * dnz;
        normalize3(n.x, n.y, n.z);
      }
      if (SCA < XXX) {  // scattering: HG with probability 1-sf, simplified Liu otherwise
        SCA = 0;
        if (COUNT) n_scat++;
        const int hk = (EXT && hole) ? 1 : 0;
        const float sf = hk ? P.SF : P.sf, g = hk ? P.G : P.g, gr = hk ? P.GR : P.gr;
        float xi = s();
        if (xi > sf) {  // divisions by constants use the reciprocals prepared on the host
          xi = (1 - xi) * P.hg_isf1[hk];
          xi = 2 * xi - 1;
          if (g != 0) {
            float ga = P.hg_a[hk] * frcp(1 + g * xi);
            xi = (ga * ga - P.hg_b[hk]) * P.hg_i2g[hk];
          }
        } else {
          xi *= P.hg_isf[hk];
          xi = 2 * (1 - fpow(xi, gr));
        }
        xi = fminf(fmaxf(xi, 0.f), 2.f);
        if (ADJ && G != 0.f && !(EXT && hole))
          atomicAdd(&gacc[P.size + J], G * w.la * P.adj_ds1db[J] * frcp(w.sca(T, J)));
        // (k1, k2, k1 k2) of the layer and their reciprocals, prepared on the host.  Where
        // every layer has k1 = k2 = 1 the frame is the identity: the way in and back (two
        // linear maps and two normalisations) would only round, and is left out
        if (P.scat_frame && !(EXT && hole)) {
          const float4 a = T.az[J], ar = T.azr[J];
          // into the frame of anisotropic scattering
          float n1 = (P.azx * n.x + P.azy * n.y) * a.x;
          float n2 = (-P.azy * n.x + P.azx * n.y) * a.y;
          n.x = n1 * P.azx - n2 * P.azy;
          n.y = n1 * P.azy + n2 * P.azx;
          n.z *= ar.z;
          normalize3(n.x, n.y, n.z);
          rotate<false>(xi, fsqrt(xi * (2 - xi)), n, s);  // the way back ends in a normalisation
          // and back
          n1 = (P.azx * n.x + P.azy * n.y) * ar.x;
          n2 = (-P.azy * n.x + P.azx * n.y) * ar.y;
          n.x = n1 * P.azx - n2 * P.azy;
          n.y = n1 * P.azy + n2 * P.azx;
          n.z *= ar.w;
          normalize3(n.x, n.y, n.z);
        } else  // the rotation keeps |n| (see rotate): no renormalisation between scatterings
          rotate<false>(xi, fsqrt(xi * (2 - xi)), n, s);
      }
    }
    ph.r = r, ph.n = n;
    ph.TOT = TOT, ph.SCA = SCA;
    ph.old = SCA == 0 ? -1 : old;  // the next step draws SCA and forgets the last DOM
    }
  }

#undef info
  if (ADJ) {  // block sums into the FP64 gradient
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * P.size; k += blockDim.x)
      if (gacc[k] != 0.f) atomicAdd(&P.adj_grad[k], (double)gacc[k]);
  } else
    P.rs[gid] = make_uint2(s.x, s.c);

  // per-launch counters: one atomic per wave; steps, photons and killed photons are
  // already wave totals, the others are summed over the lanes
  constexpr int NC = COUNT ? 8 : 4;
  unsigned long long v[8] = {n_steps, n_photons, n_killed, 0, n_cross, n_str, n_dom, n_scat};
  for (int q = 4; q < NC; q++) {
    unsigned long long t = v[q];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    v[q] = t;
  }
  if (lane == 0) {
    unsigned long long* dst[8] = {&P.ctr->steps, &P.ctr->photons, &P.ctr->killed, &P.ctr->nonfinite,
                                  &P.ctr->crossings, &P.ctr->strings, &P.ctr->doms, &P.ctr->scatters};
    for (int q = 0; q < NC; q++)
      if (v[q]) atomicAdd(dst[q], v[q]);
    atomicAdd(&P.ctr->waves, 1u);  // checked on the host: every wave of the launch ran to here
    const unsigned long long t = wall_clock64();
    atomicMax(&P.ctr->t_end_min_c, ~t);
    atomicMax(&P.ctr->t_end_max, t);
  }
}

// ---------------------------------------------------------------- launcher

typedef void (*kfun)(const KParams*);

template <bool COUNT, bool ADJ>
static kfun pick(int tilt, bool ext) {
  if (ext) {
    if (tilt == 1) return propagate<1, true, COUNT, ADJ>;
    if (tilt == 2) return propagate<2, true, COUNT, ADJ>;
    return propagate<0, true, COUNT, ADJ>;
  }
  if (tilt == 1) return propagate<1, false, COUNT, ADJ>;
  if (tilt == 2) return propagate<2, false, COUNT, ADJ>;
  return propagate<0, false, COUNT, ADJ>;
}

__global__ void reset_counters(Counters* ctr) {
  uint32_t* w = reinterpret_cast<uint32_t*>(ctr);
  for (uint32_t k = threadIdx.x; k < sizeof(Counters) / 4; k += blockDim.x) w[k] = 0;
}

// a kernel rather than a memset, so that captured graphs hold no memset nodes
hipError_t launch_reset(Counters* ctr, hipStream_t stream) {
  hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, stream, ctr);
  return hipGetLastError();
}

hipError_t launch_propagate(const KParams& P, const KParams* dP, int blocks, int threads, size_t lds_bytes,
                            hipStream_t stream, bool count, bool adjoint) {
  const bool ext = P.sum > 0 || P.hr > 0 || P.fr != 0;
  kfun f = adjoint ? pick<false, true>(P.tmod, ext)
                   : count ? pick<true, false>(P.tmod, ext) : pick<false, false>(P.tmod, ext);
  lds_bytes += sizeof(uint4) * threads;  // per-lane photon records
  if (adjoint) lds_bytes += sizeof(float) * 2 * P.size;
  hipLaunchKernelGGL(f, dim3(blocks), dim3(threads), lds_bytes, stream, dP);
  return hipGetLastError();
}

}  // namespace ppc
