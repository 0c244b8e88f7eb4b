// ref_harness.cpp -- C entry points around the reference's own CPU build.
//
// TEST INFRASTRUCTURE ONLY (see ppc_oracle.h).  This translation unit compiles
// the reference's gpu/ppc.cu from where it lies (the include path is given by
// oracle/Makefile) in its single-thread CPU form with -DXLIB -DDTMN, and adds a
// few extern "C" functions so that tests can
//   * export the tables ini::set() built (gpu/ini.cxx:578-1700) into orc_tables,
//   * run the reference kernel (gpu/pro.cu:287-904) on given bunches from a
//     given MWC state, in the single-thread order of gpu/ppc.cu:439-440, and
//     collect its raw hits.
// Nothing of the reference is copied: it is #included unchanged.
#include "ppc.cu"

#include "ppc_oracle.h"

using namespace xppc;

extern "C" {

int ref_initialize(void) {
  initialize(1.f);
  return 0;
}

// choose(dev) + ini(): one stream (seed = dev), buffers for 512*1024 photons
void ref_start(int dev) {
  choose(dev);
  ini();
}

uint32_t ref_pmax(void) { return pmax; }

float ref_zshift(float x, float y, float zz, float* dz) {
  float4 r;
  r.x = x, r.y = y, r.z = zz, r.w = 0;
  float h = d.dh;
  float s = zshift(d, r, h);
  if (dz) *dz = h;
  return s;
}

void ref_export(orc_tables* t) {
  t->size = d.size;
  t->dh = d.dh, t->rdh = d.rdh, t->hmin = d.hmin;
  t->ocv = d.ocv, t->sf = d.sf, t->g = d.g, t->gr = d.gr, t->xR = d.xR;
  t->SF = d.SF, t->G = d.G, t->GR = d.GR;
  t->hr = d.hr, t->hr2 = d.hr2, t->hs = d.hs, t->ha = d.ha, t->hifl = d.hifl;
  t->azx = d.azx, t->azy = d.azy, t->k1 = d.k1, t->k2 = d.k2, t->kz = d.kz, t->fr = d.fr;
  t->sum = d.sum;
  for (int i = 0; i < 12; i++) t->bfr[i] = d.bfr[i];
  for (int i = 0; i < MAXLYS; i++) {
    t->az_k1[i] = d.az[i].k1, t->az_k2[i] = d.az[i].k2;
    t->az_ra[i] = d.az[i].ra, t->az_rb[i] = d.az[i].rb;
  }
  t->rx = d.rx;
  for (int m = 0; m < 2; m++) t->cn[m] = d.cn[m], t->cl[m] = d.cl[m], t->crst[m] = d.crst[m];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) t->cb[i * 2 + j] = d.cb[i][j];
  for (int i = 0; i < NSTR; i++) t->ls[i] = d.ls[i];
  for (int i = 0; i < CX; i++)
    for (int j = 0; j < CY; j++) t->is_[i * CY + j] = d.is[i][j];
  // number of string lines: count entries referenced by the cell lists
  int nsc = 0;
  for (int i = 0; i < NSTR; i++) {
    const line& s = d.sc[i];
    t->sc[i].n = s.n, t->sc[i].max = s.max;
    t->sc[i].x = s.x, t->sc[i].y = s.y, t->sc[i].r = s.r;
    t->sc[i].h = s.h, t->sc[i].d = s.d, t->sc[i].dl = s.dl, t->sc[i].dh = s.dh;
  }
  for (int i = 0; i < d.cn[0]; i++)
    for (int j = 0; j < d.cn[1]; j++)
      for (unsigned short k = d.is[i][j]; k != 0x8000;) {
        unsigned short m = d.ls[k];
        if ((m & 0x7fff) + 1 > nsc) nsc = (m & 0x7fff) + 1;
        k = m & 0x8000 ? 0x8000 : k + 1;
      }
  t->nsc = nsc;
  t->tmod = d.tmod, t->vthk = d.vthk, t->lnum = d.lnum, t->lpts = d.lpts, t->l0 = d.l0;
  t->lmin = d.lmin, t->lrdz = d.lrdz, t->r0 = d.r0, t->lnx = d.lnx, t->lny = d.lny;
  for (int i = 0; i < LMAX; i++) t->lr[i] = d.lr[i];
  for (int m = 0; m < 2; m++) {
    t->mmin[m] = d.mmin[m], t->mstp[m] = d.mstp[m], t->mnum[m] = d.mnum[m], t->mcut[m] = d.mcut[m];
  }
  for (int i = 0; i < CTX; i++)
    for (int j = 0; j < CTY; j++) t->mcol[i * CTY + j] = d.mcol[i][j];
  for (int i = 0; i < LMAX; i++)
    for (int j = 0; j < LYRS; j++) {
      t->lp_x[i * LYRS + j] = z.lp[i][j].x;
      t->lp_y[i * LYRS + j] = z.lp[i][j].y;
    }
  for (int n = 0; n < WNUM; n++) {
    const ices& w = z.w[n];
    t->wvl[n] = w.wvl, t->ocm[n] = w.ocm, t->coschr[n] = w.coschr, t->sinchr[n] = w.sinchr;
    for (int i = 0; i < MAXLYS; i++) {
      t->abs_[n * MAXLYS + i] = w.z[i].abs;
      t->sca[n * MAXLYS + i] = w.z[i].sca;
    }
    t->wvs_w[n] = q.wvs[n].w, t->wvs_i[n] = q.wvs[n].i, t->wvs_f[n] = q.wvs[n].f;
  }
  t->eff = q.eff;
  t->gsize = d.gsize;
  for (int i = 0; i < d.gsize; i++) {
    const DOM& o = q.oms[i];
    t->oms[i].R = o.R, t->oms[i].F = o.F;
    t->oms[i].x = o.r[0], t->oms[i].y = o.r[1], t->oms[i].z = o.r[2];
    t->str[i] = q.names[i].str, t->dom[i] = q.names[i].dom;
  }
  t->rsize = d.rsize;
  t->l0 = d.l0;
}

uint32_t ref_multiplier(uint32_t k) { return z.rm[k]; }

static void from_photon(const photon& p, orc_bunch& b) {
  b.r[0] = p.r.x, b.r[1] = p.r.y, b.r[2] = p.r.z, b.r[3] = p.r.w;
  b.n[0] = p.n.x, b.n[1] = p.n.y, b.n[2] = p.n.z, b.n[3] = p.n.w;
  b.q = p.q, b.num = p.num, b.type = p.type, b.f = p.f;
  if (p.type > 0) {
    b.ka = p.ka, b.up = p.up, b.fldr = p.fldr, b.fla = p.fla, b.ofla = p.ofla;
  } else {
    b.a = p.a, b.b = p.b, b.beta = p.beta, b.tau = p.tau;
  }
}

// The bunch addp() builds for a particle (gpu/f2k.cxx:484-504) with direction n,
// without its Poisson draw: returns the mean photon count.
double ref_particle(float x, float y, float zz, float t, float E, float dr, int type, float nx, float ny, float nz,
                    orc_bunch* out) {
  p.n.x = nx, p.n.y = ny, p.n.z = nz;
  p.r.x = x, p.r.y = y, p.r.z = zz, p.r.w = t;
  p.type = -type;
  p.beta = 1, p.tau = 0;
  double mean = yield(E, dr, type) / ovr;
  from_photon(p, *out);
  return mean;
}

// The bunches the IceTray entry points addp_clst() and addp_mopo() build (gpu/f2k.cxx:507-528)
// for direction n (set as sett() does), taken from the reference's own global bunch after
// the call; the photons they queue are dropped.  ref_mopo returns the Poisson mean of its count.
void ref_clst(float x, float y, float zz, float t, float nx, float ny, float nz, float dr, float beta,
              orc_bunch* out) {
  p.n.x = nx, p.n.y = ny, p.n.z = nz;
  addp_clst(x, y, zz, t, 1, dr, beta);
  from_photon(p, *out);
  pn = 0, pk = 0;
}

double ref_mopo(float x, float y, float zz, float t, float nx, float ny, float nz, float light, float length,
                float beta, float tau, float fr, orc_bunch* out) {
  p.n.x = nx, p.n.y = ny, p.n.z = nz;
  addp_mopo(x, y, zz, t, light, length, beta, tau, fr);
  from_photon(p, *out);
  pn = 0, pk = 0;
  return light * q.eff / ovr;
}

// The bunch addp() builds for light of the flset()-configured flasher (type < 0: -1 a linear
// cone, -2 a 2-d gaussian, -3 isotropic, all around the direction n that sett() set, -4 the
// flasher's own beam; gpu/f2k.cxx:484-504), the components llh fits (llh/llh.cxx:1484-1492),
// taken from the reference's own global bunch after the call; the photons it queues are
// dropped.  Returns the Poisson mean of its count.
double ref_led(float x, float y, float zz, float t, float E, int type, float nx, float ny, float nz,
               orc_bunch* out) {
  p.n.x = nx, p.n.y = ny, p.n.z = nz;
  addp(x, y, zz, t, E, 0, type);
  from_photon(p, *out);
  pn = 0, pk = 0;
  return yield(E, 0, type) / ovr;
}

// per DOM, what print() reads for its acceptance (gpu/f2k.cxx:269-324): module type,
// wavelength type, relative efficiency, high voltage, cable azimuth and axis, then the
// maximum and the class efficiency of its (module, wavelength) class, whose acceptance per
// wavelength bin goes to rat (up to cap entries).  Returns that table's length.
int ref_dom_accept(int i, float* out, float* rat, int cap) {
  name& n = q.names[i];
  out[0] = n.omt, out[1] = n.type, out[2] = n.rde, out[3] = n.hv, out[4] = n.azi;
  out[5] = n.tilt[0], out[6] = n.tilt[1], out[7] = n.tilt[2];
  irde& r = irdes[std::make_pair(n.omt, n.type)];
  out[8] = r.rmax, out[9] = r.rde;
  const int m = (int)r.rat.size();
  for (int k = 0; k < m && k < cap; k++) rat[k] = r.rat[k];
  return m;
}

// the host generator (gpu/f2k.cxx:1-10) and its seed
float ref_xrnd() { return xrnd(); }
float ref_grnd() { return grnd(); }
unsigned int ref_seed() { return sv; }

// flshift(m, n, r) (gpu/f2k.cxx:590-628): n and r turned by the rotation that takes the z axis
// onto the DOM axis m (n with a NaN z: the LED elevation of the last flset() first)
void ref_flshift(float* m, float* n, float* r) { flshift(m, n, r); }

// the flasher bunch flset() builds (gpu/f2k.cxx:630-759)
void ref_flasher(int str, int dom, orc_bunch* out) {
  flset(str, dom);
  from_photon(p, *out);
}

static void to_photon(const orc_bunch& b, photon& p) {
  p.r.x = b.r[0], p.r.y = b.r[1], p.r.z = b.r[2], p.r.w = b.r[3];
  p.n.x = b.n[0], p.n.y = b.n[1], p.n.z = b.n[2], p.n.w = b.n[3];
  p.q = b.q, p.num = b.num, p.type = b.type, p.f = b.f;
  if (b.type > 0) {
    p.ka = b.ka, p.up = b.up, p.fldr = b.fldr;
    p.fla = (short)b.fla, p.ofla = (short)b.ofla;
  } else {
    p.a = b.a, p.b = b.b, p.beta = b.beta, p.tau = b.tau;
  }
}

// Run num photons of the given bunches through the reference kernel with nthr
// single-thread "threads" sharing MWC stream `stream` (multiplier mult, state
// x | c<<32).  Returns the number of raw hits (copied up to cap).
uint32_t ref_run(const orc_bunch* bunches, int nb, uint32_t num, uint32_t nthr, uint32_t stream,
                 uint32_t mult, uint32_t sx, uint32_t sc, orc_hit* hits, uint32_t cap, uint32_t* out_x,
                 uint32_t* out_c) {
  if (num > pmax || (unsigned)nb > pmxo || nthr < 1) return 0xffffffffu;
  seed = stream;
  z.rm[stream % d.rsize] = mult;
  z.rs[stream % d.rsize] = sx | (unsigned long long)sc << 32;
  for (int k = 0; k < nb; k++) to_photon(bunches[k], q.pz[k]);
  d.hidx = 0;
  d.blockIdx = 0, d.gridDim = 1, blockDim.x = nthr;
  for (threadIdx.x = 0; threadIdx.x < (int)nthr; threadIdx.x++) propagate(e, num);
  unsigned long long st = z.rs[stream % d.rsize];
  if (out_x) *out_x = (uint32_t)st;
  if (out_c) *out_c = (uint32_t)(st >> 32);
  uint32_t nh = d.hidx;
  uint32_t m = nh < d.hnum ? nh : d.hnum;
  if (m > cap) m = cap;
  for (uint32_t k = 0; k < m; k++) {
    const hit& h = q.hits[k];
    hits[k].i = h.i, hits[k].n = h.n, hits[k].z = h.z, hits[k].t = h.t;
    hits[k].pth = h.pth, hits[k].pph = h.pph, hits[k].dth = h.dth, hits[k].dph = h.dph;
  }
  return nh;
}

}  // extern "C"
