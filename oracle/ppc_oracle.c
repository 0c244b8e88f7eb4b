/*
 * ppc_oracle.c -- CPU restatement of the reference's propagation kernel.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernel (see ppc_oracle.h).
 *
 * The arithmetic follows gpu/pro.cu in its single-thread CPU form (XCPU):
 * rsqrtf is 1/sqrtf, float->int rounding is lroundf/ceilf/floorf (pro.cu:6-9),
 * and expressions keep the reference's operand order, so that a build with the
 * reference's flags (-O2 --fast-math) reproduces its results photon by photon.
 */
#include "ppc_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define FPI 3.141592653589793f
#define XXX 1.e-5f
#define OMR 0.16510f

static const float fcv = FPI / 180.f;

static inline float sq(float x) { return x * x; }
static inline float minf_(float a, float b) { return a < b ? a : b; }
/* finite test on the bits: the oracle is built with --fast-math, under which isfinite() may
 * be folded to true */
static inline int finite_(float x) {
  uint32_t u;
  memcpy(&u, &x, sizeof u);
  return (u & 0x7f800000u) != 0x7f800000u;
}
static inline float maxf_(float a, float b) { return a > b ? a : b; }
static inline int mini_(int a, int b) { return a < b ? a : b; }
static inline int maxi_(int a, int b) { return a > b ? a : b; }
static inline int rn(float x) { return (int)lroundf(x); }
static inline int ru(float x) { return (int)ceilf(x); }
static inline int rd(float x) { return (int)floorf(x); }
static inline float rsq(float x) { return 1 / sqrtf(x); }

/* ---------------------------------------------------------------- tables */

orc_tables* orc_tables_new(void) { return (orc_tables*)calloc(1, sizeof(orc_tables)); }
void orc_tables_free(orc_tables* t) { free(t); }

#define FIELD(nm)                                 \
  if (strcmp(name, #nm) == 0) {                   \
    *bytes = sizeof(t->nm);                       \
    return (void*)&t->nm;                         \
  }

void* orc_field(orc_tables* t, const char* name, uint64_t* bytes) {
  FIELD(size) FIELD(dh) FIELD(rdh) FIELD(hmin)
  FIELD(ocv) FIELD(sf) FIELD(g) FIELD(gr) FIELD(xR)
  FIELD(SF) FIELD(G) FIELD(GR) FIELD(hr) FIELD(hr2) FIELD(hs) FIELD(ha) FIELD(hifl)
  FIELD(azx) FIELD(azy) FIELD(k1) FIELD(k2) FIELD(kz) FIELD(fr) FIELD(sum) FIELD(bfr)
  FIELD(az_k1) FIELD(az_k2) FIELD(az_ra) FIELD(az_rb)
  FIELD(rx) FIELD(cn) FIELD(cl) FIELD(crst) FIELD(cb) FIELD(ls) FIELD(is_) FIELD(nsc) FIELD(sc)
  FIELD(tmod) FIELD(vthk) FIELD(lnum) FIELD(lpts) FIELD(l0)
  FIELD(lmin) FIELD(lrdz) FIELD(r0) FIELD(lnx) FIELD(lny) FIELD(lr)
  FIELD(mmin) FIELD(mstp) FIELD(mnum) FIELD(mcut) FIELD(mcol) FIELD(lp_x) FIELD(lp_y)
  FIELD(wvl) FIELD(ocm) FIELD(coschr) FIELD(sinchr) FIELD(abs_) FIELD(sca)
  FIELD(wvs_w) FIELD(wvs_i) FIELD(wvs_f) FIELD(eff)
  FIELD(gsize) FIELD(oms) FIELD(str) FIELD(dom) FIELD(rsize)
  *bytes = 0;
  return NULL;
}

/* ------------------------------------------------------ random numbers */

/* MWC step t = c + a*x (64 bit), x = lo(t), c = hi(t); uniform in (0,1]
 * from the top 23 bits of x (pro.cu:47-57) */
float orc_xrnd(orc_rng* s) {
  uint64_t t = (uint64_t)s->c + (uint64_t)s->a * (uint64_t)s->x;
  s->x = (uint32_t)t;
  s->c = (uint32_t)(t >> 32);
  union {
    uint32_t i;
    float f;
  } u;
  u.i = (s->x >> 9) | 0x3f800000u;
  return 2.f - u.f;
}

/* gamma(k) variate: Weibull rejection for k<1, Cheng's method otherwise (pro.cu:59-85) */
float orc_mrnd(float k, orc_rng* s) {
  float x;
  if (k < 1) {
    float c = 1 / k;
    float d = (1 - k) * powf(k, 1 / (c - 1));
    float zz, e;
    do {
      zz = -logf(orc_xrnd(s));
      e = -logf(orc_xrnd(s));
      x = powf(zz, c);
    } while (zz + e < d + x);
  } else {
    float b = k - logf(4.0f);
    float l = sqrtf(2 * k - 1);
    float c = 1 + logf(4.5f);
    float u, v, y, zz, r;
    do {
      u = orc_xrnd(s);
      v = orc_xrnd(s);
      y = -logf(1 / v - 1) / l;
      x = k * expf(y);
      zz = u * v * v;
      r = b + (k + l) * y - x;
    } while (r < 4.5f * zz - c && r < logf(zz));
  }
  return x;
}

/* Box-Muller (pro.cu:87-89): radius from the first draw, angle from the second */
float orc_grnd(orc_rng* s) {
  float r = sqrtf(-2 * logf(orc_xrnd(s)));
  return r * sinf(2 * FPI * orc_xrnd(s));
}

/* both Box-Muller variates (pro.cu:91-95): p[0] = r cos, p[1] = r sin */
static void grnd2(float p[2], orc_rng* s) {
  float r = sqrtf(-2 * logf(orc_xrnd(s)));
  float a = 2 * FPI * orc_xrnd(s);
  p[1] = sinf(a);
  p[0] = cosf(a);
  p[0] *= r, p[1] *= r;
}

static void norm3(float n[3]) {
  float r = rsq(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] *= r, n[1] *= r, n[2] *= r;
}

static void swapf(float* a, float* b) {
  float t = *a;
  *a = *b;
  *b = t;
}

/* Rotate unit vector n by polar angle theta (cs = 1-cos theta, si = sin theta)
 * around a uniformly drawn azimuth (pro.cu:111-153).  The largest component is
 * moved to x first so that the two perpendicular axes are well conditioned. */
void orc_rotate(float cs, float si, float n[4], orc_rng* s) {
  float p1[3], p2[3];
  int ax = 0;
  {
    float r0 = n[0] * n[0], r1 = n[1] * n[1], r2 = n[2] * n[2];
    if (r1 > r2) {
      if (r1 > r0) swapf(&n[0], &n[1]), swapf(&r0, &r1), ax = 1;
    } else {
      if (r2 > r0) swapf(&n[0], &n[2]), swapf(&r0, &r2), ax = 2;
    }
    float a = rsq(r0 + r1);
    p1[0] = -n[1] * a, p1[1] = n[0] * a, p1[2] = 0;
    float b = rsq(r0 + r2);
    p2[0] = -n[2] * b, p2[1] = 0, p2[2] = n[0] * b;
  }
  {
    float q[3];
    q[0] = p1[0] - p2[0], q[1] = p1[1] - p2[1], q[2] = p1[2] - p2[2];
    p2[0] += p1[0], p2[1] += p1[1], p2[2] += p1[2];
    float w = rsq(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    p1[0] = q[0] * w, p1[1] = q[1] * w, p1[2] = q[2] * w;
    w = rsq(p2[0] * p2[0] + p2[1] * p2[1] + p2[2] * p2[2]);
    p2[0] *= w, p2[1] *= w, p2[2] *= w;
  }
  {
    float xi = 2 * FPI * orc_xrnd(s);
    float ps = sinf(xi), pc = cosf(xi);
    n[0] -= cs * n[0] - si * (pc * p1[0] + ps * p2[0]);
    n[1] -= cs * n[1] - si * (pc * p1[1] + ps * p2[1]);
    n[2] -= cs * n[2] - si * (pc * p1[2] + ps * p2[2]);
    norm3(n);
    if (ax == 1)
      swapf(&n[0], &n[1]);
    else if (ax == 2)
      swapf(&n[0], &n[2]);
  }
}

/* 1 - cos of the scattering angle: Henyey-Greenstein with probability 1-sf,
 * simplified Liu otherwise (pro.cu:848-862) */
float orc_scatter(float sf, float g, float gr, orc_rng* s) {
  float xi = orc_xrnd(s);
  if (xi > sf) { /* Henyey-Greenstein */
    xi = (1 - xi) / (1 - sf);
    xi = 2 * xi - 1;
    if (g != 0) {
      float ga = (1 - g * g) / (1 + g * xi);
      xi = (ga * ga - sq(1 - g)) / (2 * g);
    }
  } else { /* simplified Liu */
    xi /= sf;
    xi = 2 * (1 - powf(xi, gr));
  }
  if (xi < 0)
    xi = 0;
  else if (xi > 2)
    xi = 2;
  return xi;
}

/* ------------------------------------------------------------------ tilt */

static int findcol(const orc_tables* t, int x, int y) {
  return x >= 0 && x < t->mnum[0] && y >= 0 && y < t->mnum[1] ? t->mcol[x * ORC_CTY + y] : -1;
}

/* Depth shift of the ice layers at r (pro.cu:168-264).  tmod 1: linear
 * interpolation between the reference strings along the gradient; tmod 2:
 * triangular interpolation in the skewed 2-D map with hexagon extrapolation.
 * *dz is the layer thickness, corrected when vthk is set. */
float orc_zshift(const orc_tables* t, const float r[4], float* dz) {
  if (t->tmod == 1) {
    float nr = t->lnx * r[0] + t->lny * r[1] - t->r0;
    int j = 1;
    for (; j < ORC_LMAX; j++)
      if (nr < t->lr[j] || j == t->lnum - 1) break;
    float zz = (r[2] - t->lmin) * t->lrdz;
    int i = j - 1, k = mini_(maxi_(rd(zz), 0), t->lpts);
    float pix = t->lp_x[i * ORC_LYRS + k], piy = t->lp_y[i * ORC_LYRS + k];
    float pjx = t->lp_x[j * ORC_LYRS + k], pjy = t->lp_y[j * ORC_LYRS + k];
    float dw = t->lr[j] - t->lr[i];
    float wx = (nr - t->lr[i]) / dw;
    float wy = (t->lr[j] - nr) / dw;
    float qx = pjx * wx + pix * wy;
    float qy = pjy * wx + piy * wy;
    float res = qx + qy * (zz - k);
    if (t->vthk != 0) *dz /= 1 - t->lrdz * qy;
    return res;
  } else if (t->tmod == 2) {
    float qx = t->cb[0] * r[0] + t->cb[2] * r[1];
    float qy = t->cb[1] * r[0] + t->cb[3] * r[1];
    qx = (qx - t->mmin[0]) / t->mstp[0];
    qy = (qy - t->mmin[1]) / t->mstp[1];
    {
      float x = qx, y = qy, xy;
      int n0 = t->mnum[0] - 1, n1 = t->mnum[1] - 1, c0 = t->mcut[0] - 1, c1 = t->mcut[1] - 1;
      if (y < 0)
        if (x - y > 0 && x <= c0) { x = (x - y) * c0 / (c0 - y); y = 0; }
      if (x < 0)
        if (y - x >= 0 && y < c1) { y = (y - x) * c1 / (c1 - x); x = 0; }
      if (y > n1)
        if (x >= n1 - c1 && y - x > n1 - n0) {
          x = n1 - c1 + (n0 - (n1 - c1)) * (x - (n1 - c1)) / (y - (n1 - n0) - (n1 - c1));
          y = n1;
        }
      if (x > n0)
        if (y > n0 - c0 && y - x <= n1 - n0) {
          y = n0 - c0 + (n1 - (n0 - c0)) * (y - (n0 - c0)) / (x - (n0 - n1) - (n0 - c0));
          x = n0;
        }
      if (y - x > c1)
        if (y >= c1 && x < n1 - c1) {
          xy = y - x - c1;
          xy = c1 + (x + y - c1 + xy) * (n1 - c1) / (n1 - c1 + xy);
          x = (xy - c1) / 2;
          y = (xy + c1) / 2;
        }
      if (x - y > c0)
        if (x > c0 && y <= n0 - c0) {
          xy = x - y - c0;
          xy = c0 + (x + y - c0 + xy) * (n0 - c0) / (n0 - c0 + xy);
          x = (xy + c0) / 2;
          y = (xy - c0) / 2;
        }
      qx = x, qy = y;
    }
    int nx = mini_(maxi_(rd(qx), 0), t->mnum[0]);
    int ny = mini_(maxi_(rd(qy), 0), t->mnum[1]);
    float fx = qx - nx, fy = qy - ny;
    float qrx, qry;
    int qnx, qny;
    if (fx > fy) {
      qrx = fx, qry = fy, qnx = nx + 1, qny = ny;
    } else {
      qrx = fy, qry = fx, qnx = nx, qny = ny + 1;
    }
    int cx = findcol(t, nx, ny), cy = findcol(t, qnx, qny), cz = findcol(t, nx + 1, ny + 1);
    float wx = 1 - qrx, wy = qrx - qry, wz = qry;
    float zz = (r[2] - t->lmin) * t->lrdz;
    int k = mini_(maxi_(rd(zz), 0), t->lpts);
    float ax = 0, ay = 0;
    if (cx >= 0) ax += t->lp_x[cx * ORC_LYRS + k] * wx, ay += t->lp_y[cx * ORC_LYRS + k] * wx;
    if (cy >= 0) ax += t->lp_x[cy * ORC_LYRS + k] * wy, ay += t->lp_y[cy * ORC_LYRS + k] * wy;
    if (cz >= 0) ax += t->lp_x[cz * ORC_LYRS + k] * wz, ay += t->lp_y[cz * ORC_LYRS + k] * wz;
    float res = ax + ay * (zz - k);
    if (t->vthk != 0) *dz /= 1 - t->lrdz * ay;
    return res;
  }
  return 0;
}

/* -------------------------------------------------------------- emission */

/* Emit photon idx of the launch: pick its bunch, a wavelength bin and the
 * source-specific start point and direction (pro.cu:334-477). */
void orc_emit(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t idx, orc_rng* s, orc_pbuf* out) {
  int pj = -1;
  uint64_t pn = 0;
  while (idx >= pn && pj + 1 < nb) pn += pz[++pj].num;
  const orc_bunch* p = &pz[pj];

  int j = rd(ORC_WNUM * orc_xrnd(s));
  if (j > ORC_WNUM - 1) j = ORC_WNUM - 1;
  float r[4] = {p->r[0], p->r[1], p->r[2], p->r[3]};
  float n[4] = {p->n[0], p->n[1], p->n[2], t->ocm[j]};
  float l = p->n[3];
  int fla, ofla;

  if (p->type > 0) { /* flasher / LED */
    fla = p->fla, ofla = p->ofla;
    if (l == 0) { /* azimuthally symmetric or N-fold LED pattern */
      float xi = orc_xrnd(s);
      if (p->fldr < 0)
        xi *= 2 * FPI;
      else {
        int rr = rd(p->fldr / 360) + 1;
        int ss = rd(xi * rr);
        xi = (p->fldr + ss * 360 / rr) * fcv;
      }
      r[1] = sinf(xi), r[0] = cosf(xi);
      float cu = cosf(p->up);
      n[2] = sinf(p->up);
      n[0] = cu * r[0];
      n[1] = cu * r[1];
    }
    if (p->ka != 0) {
      float xi;
      if (p->ka > 999.f) { /* lab-measured beam profile */
        xi = orc_xrnd(s);
        float s1 = 0.0191329f, s2 = 0.0686944f, C1 = 0.583583f, C2 = 0.967762f;
        if (p->type == 1) s1 = 0.0196242f, s2 = 0.0750278f, C1 = 0.606833f, C2 = 0.960863f;
        if (p->type == 2) s1 = 0.0185603f, s2 = 0.0624481f, C1 = 0.553192f, C2 = 0.974129f;
        if (xi < C1)
          xi = s1 * fabsf(orc_grnd(s));
        else if (xi < C2)
          xi = -s2 * logf(orc_xrnd(s));
        else
          xi = 2.f;
        if (xi >= 2.f) xi = 2 * (1 - sqrtf(orc_xrnd(s)));
      } else if (p->ka > 0) /* 2-d gaussian */
        xi = minf_(-p->ka * logf(1 - (1 - orc_xrnd(s)) * (1 - expf(-2 / p->ka))), 2.f);
      else if (p->ka == -1) /* isotropic */
        xi = 2 * (1 - orc_xrnd(s));
      else /* linear cone */
        xi = 2 * (1 - sqrtf(orc_xrnd(s)));
      orc_rotate(xi, sqrtf(xi * (2 - xi)), n, s);
    }
    if (p->ka > 999.f) { /* start on the flasher board inside the DOM */
      const float FLZ = 0.07735f, FLR = 0.119f - 0.008f, FLB = 0.03396f;
      if (l == 0) r[0] *= FLR, r[1] *= FLR, r[2] = FLZ, r[3] = 0;
      float b = r[0] * n[0] + r[1] * n[1] + r[2] * n[2];
      float c = FLZ * FLZ + FLR * FLR - OMR * OMR;
      float tt = sqrtf(b * b - c) - b;
      r[0] += tt * n[0], r[1] += tt * n[1], r[2] += tt * n[2], r[3] += tt * t->ocv;
      if (fabsf(r[2]) < FLB)
        ofla = -2;
      else if (l == 0 && r[2] < 0 && orc_xrnd(s) < 0.9f)
        ofla = -3;
      if (p->ka > 1050.f) { /* cable shadow */
        const float rc = 0.023f, dc = OMR + rc;
        float ph = (p->ka - 1080.f) * fcv;
        float drx = r[0] - dc * cosf(ph);
        float dry = r[1] - dc * sinf(ph);
        float a = n[0] * n[0] + n[1] * n[1];
        if (a > 0) {
          float b2 = n[0] * drx + n[1] * dry;
          float c2 = drx * drx + dry * dry - rc * rc;
          float D = b2 * b2 - a * c2;
          if (D >= 0) {
            float h1 = (-b2 + sqrtf(D)) / a;
            if (h1 > 0) ofla = -4;
          }
        }
      }
      if (l == 0)
        r[0] += p->r[0], r[1] += p->r[1], r[2] += p->r[2], r[3] += p->r[3];
      else {
        const orc_dom* dm = &t->oms[fla];
        r[0] += dm->x, r[1] += dm->y, r[2] += dm->z;
        r[3] += p->r[3];
      }
    } else if (fla >= 0) { /* start on the DOM surface */
      const float tt = OMR;
      r[0] = tt * n[0], r[1] = tt * n[1], r[2] = tt * n[2], r[3] = tt * t->ocv;
      r[0] += p->r[0], r[1] += p->r[1], r[2] += p->r[2], r[3] += p->r[3];
    } else {
      r[0] = p->r[0], r[1] = p->r[1], r[2] = p->r[2], r[3] = p->r[3];
    }
  } else { /* particles: muon segments and cascades */
    fla = -1, ofla = -1;
    if (l > 0)
      l *= orc_xrnd(s);
    else
      l = p->b * orc_mrnd(p->a, s);
    if (l > 0) {
      r[0] += l * n[0], r[1] += l * n[1], r[2] += l * n[2];
      r[3] += l * t->ocv / p->beta;
    }
    if (p->tau > 0) { /* isotropic delayed emission */
      r[3] -= p->tau * logf(orc_xrnd(s));
      float xi = 1 - orc_xrnd(s);
      orc_rotate(xi, sqrtf(xi * (2 - xi)), n, s);
    } else {
      float cs = t->coschr[j], si = t->sinchr[j];
      if (p->f < orc_xrnd(s)) { /* cascade particle directions */
        const float a = 0.39f, b = 2.61f;
        const float I = 1 - expf(-b * exp2f(a));
        float xi = minf_(powf(-logf(1 - orc_xrnd(s) * I) / b, 1 / a), 2.f);
        orc_rotate(xi, sqrtf(xi * (2 - xi)), n, s);
      } else {
        float beta = p->beta;
        if (p->type < 0) { /* muon end point */
          const float ar = 0.26f * 0.9216f / 0.105658389f;
          float dx = ar * (p->n[3] - l);
          beta = sqrtf(dx * (2 + dx)) / (1 + dx);
          if (beta >= cs) r[3] -= t->ocv * (sqrtf(dx * dx + 1) - asinhf(1 / dx) - dx) / ar;
        }
        if (beta < 1) {
          float xi = cs / beta;
          if (xi <= 1) {
            float sx = sqrtf(1 - xi * xi);
            if (p->type < 0 && sx < orc_xrnd(s) * si) ofla = -3;
            cs = xi, si = sx;
          } else
            ofla = -2;
        }
      }
      orc_rotate(1 - cs, si, n, s); /* Cherenkov cone */
    }
  }
  memcpy(out->r, r, sizeof r);
  memcpy(out->n, n, sizeof n);
  out->q = (uint32_t)j;
  out->i = p->q;
  out->fla = fla, out->ofla = ofla;
}

/* ------------------------------------------------------------- transport */

#define ABS(t, w, J) ((t)->abs_[(w) * ORC_MAXLYS + (J)])
#define SCA(t, w, J) ((t)->sca[(w) * ORC_MAXLYS + (J)])

/* Path length of the straight segment from layer coordinate z0 (in units of
 * the layer thickness) to z0 + dz that lies in layer k: the overlap of the
 * segment with [k-1/2, k+1/2], the outermost layers reaching to infinity, scaled
 * from layer coordinates to path length (len/|dz|).  Horizontal segments lie in
 * layer round(z0).  Independent of the layer walk of the transport. */
static double layer_overlap(int k, int size, double z0, double dz, double len) {
  if (dz == 0) return k == mini_(maxi_(rn((float)z0), 0), size - 1) ? len : 0;
  double lo = z0 < z0 + dz ? z0 : z0 + dz, hi = z0 < z0 + dz ? z0 + dz : z0;
  double a = k == 0 ? -1e30 : k - 0.5, b = k == size - 1 ? 1e30 : k + 0.5;
  double o = (hi < b ? hi : b) - (lo > a ? lo : a);
  return o > 0 ? o * len / (hi - lo) : 0;
}

/* Propagate one emitted photon until it is absorbed, leaves through a DOM
 * (xR=1) or is killed, recording raw hits (pro.cu:482-889).  With A, also
 * accumulate the score-function gradient of the LLH weight of the photon's hits
 * (orc_adjoint_photon). */
static int transport(const orc_tables* t, const orc_pbuf* f, orc_rng* s, orc_hit* hits, int cap,
                     orc_counters* cnt, orc_adjoint* A);

int orc_transport(const orc_tables* t, const orc_pbuf* f, orc_rng* s, orc_hit* hits, int cap,
                  orc_counters* cnt) {
  return transport(t, f, s, hits, cap, cnt, NULL);
}

static int transport(const orc_tables* t, const orc_pbuf* f, orc_rng* s, orc_hit* hits, int cap,
                     orc_counters* cnt, orc_adjoint* A) {
  float r[4], n[4];
  memcpy(r, f->r, sizeof r);
  memcpy(n, f->n, sizeof n);
  const int w = (int)f->q;
  const uint32_t niw = f->i;
  const int ofla = f->ofla;
  int om = f->fla, old = -1;
  int nhit = 0;
  float sqd = 0;

  float TOT = -logf(orc_xrnd(s)), SCA_ = 0;
  if (ofla < -1) TOT = 0;
  if (cnt) cnt->photons++;
  if (TOT == 0 && cnt) cnt->killed++;

  for (int first = 1;; first = 0) {
    if (!first) om = -1;
    if (SCA_ == 0) SCA_ = -logf(orc_xrnd(s)), old = om;
    if (cnt) cnt->steps++;
    float sca, tot;

    int sign = n[2] < 0 ? -1 : 1;
    float anz = sign * n[2];
    float edh = t->dh;
    float zz = r[2] - orc_zshift(t, r, &edh);
    float nr = 1.f;
    int I, J;
    {
      zz = (zz - t->hmin) * t->rdh;
      I = mini_(maxi_(rn(zz), 0), t->size - 1);
      if (TOT > 0) { /* absorption anisotropy */
        float n1 = t->azx * n[0] + t->azy * n[1];
        float n2 = -t->azy * n[0] + t->azx * n[1];
        float n3 = n[2];
        if (t->fr != 0) {
          float k1 = t->az_k1[I], k2 = t->az_k2[I];
          float kz = 1 / (k1 * k2);
          float s1 = n1 * n1, l1 = k1 * k1;
          float s2 = n2 * n2, l2 = k2 * k2;
          float s3 = n3 * n3, l3 = kz * kz;
          float B2 = nr / l1 + nr / l2 + nr / l3;
          float nB = s1 / l1 + s2 / l2 + s3 / l3;
          float An = s1 * l1 + s2 * l2 + s3 * l3;
          nr = powf((B2 - nB) * An / 2, t->fr);
        }
        n1 *= t->k1, n2 *= t->k2, n3 *= t->kz;
        nr *= rsq(n1 * n1 + n2 * n2 + n3 * n3);
        TOT /= nr;
      }
      float ahx = edh * (0.5f - sign * (zz - I));
      J = I;
      sca = 0, tot = 0;
      float ais = SCA_, aia = TOT;
      for (;;) { /* walk through layers until the scattering or absorption depth is spent */
        float dsca = ais / SCA(t, w, J), dtot = aia / ABS(t, w, J);
        float dist = minf_(dsca, dtot) * anz;
        if ((sign < 0 ? J > 0 : J < t->size - 1) && dist > ahx) {
          float del = ahx / anz;
          sca += del, tot += del;
          ais -= SCA(t, w, J) * del, aia -= ABS(t, w, J) * del;
          ahx = edh;
          J += sign;
          if (cnt) cnt->crossings++;
        } else {
          sca += dsca, tot += dtot;
          break;
        }
      }
      if (sca < 0) sca = 0;
      if (tot < 0) tot = 0;
      if (tot < sca)
        sca = tot, tot = 0;
      else
        tot = (tot - sca) * ABS(t, w, J);
    }

    om = -1;
    float del = sca;
    float hi = sca, hf = 0;
    { /* DOM and hole-ice search through the cell grid */
      float rix = r[0], rfx = r[0] + sca * n[0];
      float riy = r[1], rfy = r[1] + sca * n[1];
      float pix = t->cb[0] * rix + t->cb[2] * riy, piy = t->cb[1] * rix + t->cb[3] * riy;
      float pfx = t->cb[0] * rfx + t->cb[2] * rfy, pfy = t->cb[1] * rfx + t->cb[3] * rfy;
      float lox = minf_(pix, pfx) - t->rx, hix = maxf_(pix, pfx) + t->rx;
      float loy = minf_(piy, pfy) - t->rx, hiy = maxf_(piy, pfy) + t->rx;
      int xlx = mini_(maxi_(rn((lox - t->cl[0]) * t->crst[0]), 0), t->cn[0]);
      int xhx = maxi_(mini_(rn((hix - t->cl[0]) * t->crst[0]), t->cn[0] - 1), -1);
      int xly = mini_(maxi_(rn((loy - t->cl[1]) * t->crst[1]), 0), t->cn[1]);
      int xhy = maxi_(mini_(rn((hiy - t->cl[1]) * t->crst[1]), t->cn[1] - 1), -1);
      for (int ci = xlx; ci <= xhx; ci++)
        for (int cj = xly; cj <= xhy; cj++) {
          if (cnt) cnt->cells++;
          for (unsigned k = t->is_[ci * ORC_CY + cj]; k != 0x8000;) {
            unsigned m = t->ls[k];
            const orc_line* ln = &t->sc[m & 0x7fff];
            k = m & 0x8000 ? 0x8000 : k + 1;
            if (cnt) cnt->strings++;
            float b = 0, c = 0, dr;
            dr = ln->x - r[0];
            b += n[0] * dr;
            c += dr * dr;
            dr = ln->y - r[1];
            b += n[1] * dr;
            c += dr * dr;
            float np = 1 - n[2] * n[2];
            float D = b * b - (c - ln->r * ln->r) * np;
            if (D >= 0) {
              D = sqrtf(D);
              float h1 = b - D, h2 = b + D;
              if (h2 >= 0 && h1 <= sca * np) {
                if (np > XXX) {
                  h1 /= np, h2 /= np;
                  if (h1 < 0) h1 = 0;
                  if (h2 > sca) h2 = sca;
                } else
                  h1 = 0, h2 = sca;
                h1 = r[2] + n[2] * h1, h2 = r[2] + n[2] * h2;
                float zl, zh;
                if (n[2] > 0)
                  zl = h1, zh = h2;
                else
                  zl = h2, zh = h1;
                int omin = 0, omax = ln->max;
                int n1 = ln->n - omin + mini_(omax + 1, maxi_(omin, ru(omin - (zh - ln->dl - ln->h) * ln->d)));
                int n2 = ln->n - omin + maxi_(omin - 1, mini_(omax, rd(omin - (zl - ln->dh - ln->h) * ln->d)));
                for (int l = n1; l <= n2; l++)
                  if (l != old) {
                    const orc_dom* dm = &t->oms[l];
                    if (cnt) cnt->doms++;
                    float a = 0, bb = 0, cc = 0, d2;
                    float ff = dm->F > 0 ? sq(1 / dm->F) : 0;
                    d2 = dm->x - r[0];
                    a += n[0] * n[0], bb += n[0] * d2, cc += d2 * d2;
                    d2 = dm->y - r[1];
                    a += n[1] * n[1], bb += n[1] * d2, cc += d2 * d2;
                    d2 = dm->z - r[2];
                    a += ff * n[2] * n[2], bb += ff * n[2] * d2, cc += ff * d2 * d2;
                    if (a > 0) {
                      bb /= a;
                      float xR = (l == ofla ? 1 : t->xR);
                      float DD = bb * bb - (cc - sq(xR * dm->R)) / a;
                      if (DD >= 0) {
                        sqd = sqrtf(DD);
                        float h = bb - sqd;
                        if (dm->F > 0 || fabsf(d2 - n[2] * h) < -xR * dm->F * dm->R) {
                          sqd /= xR;
                          h = bb - sqd;
                          if (h > 0 && h <= del) om = l, del = h;
                        }
                      }
                    }
                  }
              }
            }
            if (t->hr > 0 &&
                (t->hifl == 0 || (t->hifl < 0 && r[3] < -t->hifl) || (t->hifl > 0 && r[3] > t->hifl))) {
              float D2 = b * b - (c - t->hr2) * np;
              if (D2 > 0) {
                D2 = sqrtf(D2);
                float h1 = b - D2, h2 = b + D2;
                if (h2 >= 0 && h1 <= sca * np) {
                  if (np > XXX) {
                    h1 /= np, h2 /= np;
                    if (h1 < 0) h1 = 0;
                    if (h2 > sca) h2 = sca;
                  } else
                    h1 = 0, h2 = sca;
                  if (h1 < hi && h2 > sqrtf(XXX) * t->hr) hi = h1, hf = h2;
                }
              }
            }
          }
        }
    }

    float fin = minf_(del, hi);
    int hole = fin < sca;
    if (hole) { /* step cut short by a DOM or the hole-ice column */
      float xs = 0, xa = 0;
      float y = zz + n[2] * fin / edh;
      J = mini_(maxi_(rn(y), 0), t->size - 1);
      if (I == J)
        xs = fin * SCA(t, w, I), xa = fin * ABS(t, w, I);
      else {
        float h = 0.5f - sign * (zz - I);
        float g = 0.5f + sign * (y - J);
        xs = h * SCA(t, w, I) + g * SCA(t, w, J);
        xa = h * ABS(t, w, I) + g * ABS(t, w, J);
        for (int i = I + sign; i != J; i += sign) {
          xs += SCA(t, w, i);
          xa += ABS(t, w, i);
        }
        float fct = edh / anz;
        xs *= fct, xa *= fct;
      }
      if (xs < 0) xs = 0;
      if (xa < 0) xa = 0;
      SCA_ -= xs, TOT -= xa;
      TOT *= nr;
      if (hi < del) {
        float fin2 = minf_(hi + minf_(SCA_ / t->hs, TOT / t->ha), hf);
        if (om != -1) {
          if (fin2 < del) del = fin2, om = -1;
        } else {
          if (fin2 < del)
            del = fin2;
          else
            del = minf_(fin2, sca);
        }
        fin2 = del - hi;
        SCA_ -= fin2 * t->hs, TOT -= fin2 * t->ha;
      }
    } else
      SCA_ = 0, TOT = tot * nr;

    if (A && A->G != 0) { /* path-length scores of the segment, layer by layer */
      double dz = (double)n[2] * del / edh;
      for (int k = 0; k < t->size; k++) {
        double l = layer_overlap(k, t->size, zz, dz, del);
        if (l > 0) {
          A->ga[k] -= A->G * nr * A->lk[w] * A->dpda[k] * l;
          A->gb[k] -= A->G * A->la[w] * A->ds1db[k] * l;
        }
      }
    }

    r[0] += del * n[0];
    r[1] += del * n[1];
    r[2] += del * n[2];
    r[3] += del * n[3];

    /* the GPU build's guard against non-finite optical depths (pro.cu:727-729; the
     * single-thread build has none, a NaN there never ends the photon) */
    if (!finite_(TOT) || !finite_(SCA_)) {
      TOT = 0, om = -1;
      if (cnt) cnt->nonfinite++;
    }

    if (om != -1) { /* DOM hit */
      if (om == ofla)
        TOT = 0;
      else {
        orc_hit h;
        h.i = (uint32_t)om;
        h.t = r[3];
        h.n = niw;
        h.z = (uint32_t)t->wvl[w];
        h.pth = acosf(n[2]);
        h.pph = atan2f(n[1], n[0]);
        sqd *= t->xR - 1;
        const orc_dom* dm = &t->oms[om];
        float dx = dm->x - r[0] + sqd * n[0];
        float dy = dm->y - r[1] + sqd * n[1];
        float dzz = dm->z - r[2] + sqd * n[2];
        h.dth = acosf(minf_(maxf_(dzz / (t->xR * dm->R * dm->F), -1.f), 1.f));
        h.dph = atan2f(dy, dx);
        if (nhit < cap) hits[nhit] = h;
        nhit++;
        if (A && niw < A->nseg) A->G -= A->weight[(size_t)niw * A->gsize + om];
        if (cnt) cnt->hits++;
        if (t->xR == 1)
          TOT = 0;
        else
          old = om;
      }
    }

    if (TOT < XXX) {
      TOT = 0;
      break;
    }
    if (t->sum > 0 && fin > 0) { /* birefringence */
      float ra, rb;
      if (I == J)
        ra = fin * t->az_ra[I], rb = fin * t->az_rb[I];
      else {
        float y = zz + n[2] * fin / edh;
        float h = 0.5f - sign * (zz - I);
        float g = 0.5f + sign * (y - J);
        ra = h * t->az_ra[I] + g * t->az_ra[J];
        rb = h * t->az_rb[I] + g * t->az_rb[J];
        for (int i = I + sign; i != J; i += sign) {
          ra += t->az_ra[i];
          rb += t->az_rb[i];
        }
        float fct = edh / anz;
        ra *= fct, rb *= fct;
      }
      if (ra < 0) ra = 0;
      if (rb < 0) rb = 0;
      float dot = t->azx * n[0] + t->azy * n[1];
      float sdt = sqrtf(1 - dot * dot);
      float cdt = fabsf(dot);
      const float* B = t->bfr;
      float sx = sqrtf(ra) * maxf_(0.f, B[0] * expf(-B[1] * powf(atanf(B[3] * sdt), B[2])));
      float sy = sqrtf(ra) * maxf_(0.f, B[4] * expf(-B[5] * powf(atanf(B[7] * sdt), B[6])));
      float mx = rb * maxf_(0.f, B[8] * atanf(B[11] * sdt * cdt) * expf(-B[9] * sdt + B[10] * cdt));
      float qx = t->azx - dot * n[0], qy = t->azy - dot * n[1], qz = -dot * n[2];
      if (dot < 0) qx = -qx, qy = -qy, qz = -qz;
      float qn = qx * qx + qy * qy + qz * qz;
      if (qn > 0) {
        float fq = rsq(qn);
        qx *= fq, qy *= fq, qz *= fq;
      } else
        qx = 0, qy = 0, qz = 1;
      float px = n[1] * qz - n[2] * qy, py = n[2] * qx - n[0] * qz, pzz = n[0] * qy - n[1] * qx;
      float g2[2];
      grnd2(g2, s);
      float dnx = sx * g2[0] + mx;
      float dny = sy * g2[1];
      if (del <= hi) { /* curved path within the step */
        grnd2(g2, s);
        float vnx = dnx + sx * g2[0] * rsq(3.f);
        float vny = dny + sy * g2[1] * rsq(3.f);
        vnx /= 2, vny /= 2;
        float den = vnx * vnx + vny * vny;
        float dnz = -den;
        den = 2 / (1 + den);
        vnx *= den, vny *= den, dnz *= den;
        float m[3];
        m[0] = n[0] + qx * vnx + px * vny + n[0] * dnz;
        m[1] = n[1] + qy * vnx + py * vny + n[1] * dnz;
        m[2] = n[2] + qz * vnx + pzz * vny + n[2] * dnz;
        norm3(m);
        del -= fin / 2;
        r[0] += del * (m[0] - n[0]);
        r[1] += del * (m[1] - n[1]);
        r[2] += del * (m[2] - n[2]);
      }
      dnx /= 2, dny /= 2;
      float den = dnx * dnx + dny * dny;
      float dnz = -den;
      den = 2 / (1 + den);
      dnx *= den, dny *= den, dnz *= den;
      n[0] += qx * dnx + px * dny + n[0] * dnz;
      n[1] += qy * dnx + py * dny + n[1] * dnz;
      n[2] += qz * dnx + pzz * dny + n[2] * dnz;
      norm3(n);
    }
    if (SCA_ < XXX) { /* scattering */
      SCA_ = 0;
      if (cnt) cnt->scatters++;
      float sf, g, gr;
      if (hole)
        sf = t->SF, g = t->G, gr = t->GR;
      else
        sf = t->sf, g = t->g, gr = t->gr;
      if (A && A->G != 0 && !hole) A->gb[J] += A->G * A->la[w] * A->ds1db[J] / SCA(t, w, J);
      float xi = orc_scatter(sf, g, gr, s);
      float k1 = t->az_k1[J], k2 = t->az_k2[J];
      float okz = k1 * k2;
      if (!hole) { /* into the frame of anisotropic scattering */
        float n1 = (t->azx * n[0] + t->azy * n[1]) * k1;
        float n2 = (-t->azy * n[0] + t->azx * n[1]) * k2;
        n[0] = n1 * t->azx - n2 * t->azy;
        n[1] = n1 * t->azy + n2 * t->azx;
        n[2] /= okz;
        norm3(n);
      }
      orc_rotate(xi, sqrtf(xi * (2 - xi)), n, s);
      if (!hole) { /* and back */
        float n1 = (t->azx * n[0] + t->azy * n[1]) / k1;
        float n2 = (-t->azy * n[0] + t->azx * n[1]) / k2;
        n[0] = n1 * t->azx - n2 * t->azy;
        n[1] = n1 * t->azy + n2 * t->azx;
        n[2] *= okz;
        norm3(n);
      }
    }
  }
  return nhit;
}

int orc_photon(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t idx, orc_rng* s, orc_hit* hits,
               int cap, orc_counters* cnt) {
  orc_pbuf f;
  orc_emit(t, pz, nb, idx, s, &f);
  return orc_transport(t, &f, s, hits, cap, cnt);
}

int orc_adjoint_photon(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t idx, const orc_rng* s0,
                       orc_adjoint* A) {
  orc_hit h[64];
  orc_rng s = *s0;
  orc_pbuf f;
  orc_emit(t, pz, nb, idx, &s, &f);
  int nh = transport(t, &f, &s, h, 64, NULL, NULL);
  double G = 0;
  for (int k = 0; k < nh && k < 64; k++)
    if (h[k].n < A->nseg) G += A->weight[(size_t)h[k].n * A->gsize + h[k].i];
  A->G = G;
  if (G == 0) return nh;
  s = *s0;
  orc_emit(t, pz, nb, idx, &s, &f);
  transport(t, &f, &s, h, 64, NULL, A);
  return nh;
}

/* The score of a photon's complete path: the same accumulation as the second run of
 * orc_adjoint_photon with the weight held at 1 from emission to absorption (hits do not
 * lower it).  Over complete paths the score has mean zero (the paths' density integrates
 * to 1 for any ice), which pins the scattering-event term against the path-length term. */
int orc_adjoint_path(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t idx, const orc_rng* s0,
                     orc_adjoint* A, orc_counters* cnt, uint32_t* row) {
  orc_hit h[64];
  orc_rng s = *s0;
  orc_pbuf f;
  orc_emit(t, pz, nb, idx, &s, &f);
  if (row) *row = f.q;
  const uint32_t nseg = A->nseg;
  A->nseg = 0;
  A->G = 1;
  int nh = transport(t, &f, &s, h, 64, cnt, A);
  A->nseg = nseg;
  return nh;
}

uint32_t orc_run_xcpu(const orc_tables* t, const orc_bunch* pz, int nb, uint32_t num, uint32_t nthr,
                      orc_rng* s, orc_hit* hits, uint32_t cap, orc_counters* cnt) {
  uint32_t nh = 0;
  orc_pbuf* buf = (orc_pbuf*)malloc(sizeof(orc_pbuf) * (num / nthr + 1));
  for (uint32_t th = 0; th < nthr; th++) {
    uint32_t k = 0;
    for (uint32_t i = th; i < num; i += nthr) orc_emit(t, pz, nb, i, s, &buf[k++]);
    k = 0;
    for (uint32_t i = th; i < num; i += nthr) {
      int room = nh < cap ? (int)(cap - nh) : 0;
      int got = orc_transport(t, &buf[k++], s, hits + (nh < cap ? nh : 0), room, cnt);
      nh += (uint32_t)got;
    }
  }
  free(buf);
  return nh;
}
