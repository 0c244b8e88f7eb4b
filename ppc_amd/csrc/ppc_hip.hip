// ppc_hip.hip -- host side of the C-ABI in include/ppc_hip.h.
//
// Owns the device buffers of one GPU, packs the tables into the kernel's
// parameter block and LDS image, and runs launches on one HIP stream with
// events around the kernel (the roles of struct gpu in gpu/ppc.cu:102-321).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kparams.h"
#include "ppc_hip.h"

namespace ppc {
hipError_t launch_propagate(const KParams& P, const KParams* dP, int blocks, int threads, size_t lds_bytes,
                            hipStream_t stream, bool count, bool adjoint);
hipError_t launch_reset(Counters* ctr, hipStream_t stream);
hipError_t launch_accept(const AcceptParams& A, const ppc_hip_hit* hits, const uint32_t* photon, uint32_t n,
                         uint64_t seed, ppc_hip_accepted* out, hipStream_t stream);
}

using ppc::KParams;

namespace {

thread_local std::string g_err;

// clearance map entries are capped at this distance (m); the kernel's map has at most
// this many entries (16 MiB)
constexpr double kClearCap = 100.0;
constexpr size_t kClearMaxEntries = size_t(1) << 22;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) {                                                               \
      return fail(e_ == hipErrorOutOfMemory ? PPC_HIP_E_NOMEM : PPC_HIP_E_DEVICE, "%s: %s (%s:%d)", #call, \
                  hipGetErrorString(e_), __FILE__, __LINE__);                             \
    }                                                                                     \
  } while (0)

template <class T>
hipError_t dalloc(T** p, size_t n) {
  return hipMalloc((void**)p, sizeof(T) * (n > 0 ? n : 1));
}

// ---- the scratch area of ppc_hip_adjoint_noalloc ----
// One per GPU, as a file-scope device array, so that the entry point allocates nothing.
// Host code never passes the array's name to a kernel or a copy (on the host it is only a
// stand-in for the symbol): ppc_hip_open looks up its device address once per GPU with
// hipGetSymbolAddress and every later use goes through that pointer.  Calls on one GPU
// share the area: overlapping calls (two contexts on one GPU, two host threads) race.
constexpr uint32_t kScratchWeights = 1u << 18;  // LLH weights: segments x DOMs <= 262144 (1 MiB)
constexpr uint32_t kScratchRecords = 1u << 16;  // replay records the caller passes (1 MiB)
struct AdjScratch {
  double grad[2 * PPC_HIP_MAXLYS];
  float dp[2 * PPC_HIP_MAXLYS];
  ppc::Counters ctr;
  float weight[kScratchWeights];
  ppc_hip_adj_record rec[kScratchRecords];
};
static __device__ __align__(256) unsigned char scratch[sizeof(AdjScratch)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
AdjScratch* g_scratch[kMaxDevices] = {};

// device address of this GPU's scratch area (the current device must be `device`)
hipError_t scratch_address(int device, AdjScratch** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<AdjScratch*>(ptr);
  }
  *out = g_scratch[device];
  return hipSuccess;
}

__global__ void zero_words(uint32_t* p, uint32_t n) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}

hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_words, dim3(std::min<uint32_t>((n + 255) / 256, 1024)), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace

struct ppc_hip_ctx {
  int device = 0;
  int cus = 0;
  int wall_khz = 0;  // rate of the device's wall clock (wall_clock64)
  ppc_hip_config cfg{};
  int lanes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_copy = nullptr;
  hipEvent_t ev_a0 = nullptr, ev_a1 = nullptr;  // adjoint kernel
  bool launched = false, uploaded = false, copy_pending = false;

  // launch buffers
  ppc_hip_bunch* d_pz = nullptr;
  uint32_t* d_pstart = nullptr;
  ppc_hip_hit* d_hits = nullptr;
  uint32_t* d_hit_photon = nullptr;
  ppc::Counters* d_ctr = nullptr;
  ppc_hip_record* d_rec = nullptr;
  uint2* d_rs = nullptr;
  uint32_t* d_rm = nullptr;
  // pinned staging
  ppc_hip_bunch* h_pz = nullptr;
  uint32_t* h_pstart = nullptr;
  ppc::Counters* h_ctr = nullptr;
  KParams* h_P = nullptr;  // pinned staging of the launch parameters
  KParams* d_P = nullptr;  // the kernel reads its parameters from here

  // tables
  uint4* d_image = nullptr;
  float2* d_lp = nullptr;
  ppc_hip_wv* d_wv = nullptr;
  float4* d_dom = nullptr;
  float* d_domF = nullptr;
  size_t lds_bytes = 0;
  KParams P{};
  uint32_t last_num = 0;
  std::vector<uint8_t> img_base;  // LDS image without the layer table
  int nvar = 1;                   // ice variants in the layer table
  uint16_t* d_bvar = nullptr;     // per-bunch ice variant of a batch launch
  uint16_t* h_bvar = nullptr;
  float* d_clear = nullptr;       // clearance map of the string cylinders
  // hit post-processing tables and output (ppc_hip_upload_accept / ppc_hip_accept)
  ppc_hip_sensor* d_sensors = nullptr;
  float* d_pmt_dirs = nullptr;
  float* d_rat = nullptr;
  ppc_hip_dom_accept* d_adoms = nullptr;
  float* d_awvs = nullptr;
  int32_t accept_gsize = -1;
  ppc_hip_accepted* d_accepted = nullptr;
  ppc_hip_adj_record* d_hphot = nullptr;  // photons that hit (record_starts)
  uint32_t fwd_nphot = 0;         // their number in the last waited-for launch
  bool fwd_waited = false;
  ppc::Counters* d_ctr_adj = nullptr;  // the adjoint's own counters
  ppc::Counters* h_ctr_adj = nullptr;
  // adjoint
  ppc_hip_adj_record* d_adj_rec = nullptr;
  uint32_t adj_cap = 0;
  float* d_adj_w = nullptr;
  size_t adj_wcap = 0;
  float* d_adj_dp = nullptr;
  double* d_adj_grad = nullptr;
  AdjScratch* scratch = nullptr;  // this GPU's scratch area (ppc_hip_adjoint_noalloc)
  uint64_t photon_seed = 0;       // ppc_hip_seed_photons
  hipGraph_t graph = nullptr;  // ppc_hip_graph_capture: the graph and the buffers it reads
  hipGraphExec_t graph_exec = nullptr;
  uint32_t graph_num = 0;
  ppc_hip_bunch* g_pz = nullptr;
  uint32_t* g_pstart = nullptr;
  uint16_t* g_bvar = nullptr;
  KParams* g_P = nullptr;
};

namespace {
void drop_graph(ppc_hip_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  for (void* p : {(void*)c->g_pz, (void*)c->g_pstart, (void*)c->g_bvar, (void*)c->g_P})
    if (p) (void)hipFree(p);
  c->graph_exec = nullptr, c->graph = nullptr;
  c->g_pz = nullptr, c->g_pstart = nullptr, c->g_bvar = nullptr, c->g_P = nullptr;
}

// append the layer table of nvar ice variants ([nvar][size]) to the LDS image and upload it
int set_layers(ppc_hip_ctx* c, const ppc_hip_layer* lay, int nvar) {
  KParams& P = c->P;
  std::vector<uint8_t> img = c->img_base;
  P.off_lay = (int)img.size();
  img.resize(img.size() + sizeof(ppc_hip_layer) * P.size * nvar);
  memcpy(img.data() + P.off_lay, lay, sizeof(ppc_hip_layer) * P.size * nvar);
  img.resize((img.size() + 15) & ~size_t(15));
  // the image, a 16-byte photon record per lane and the adjoint's block sums share the LDS
  if (img.size() + 16 * (size_t)c->cfg.threads + sizeof(float) * 2 * P.size > 64 * 1024)
    return fail(PPC_HIP_E_CAPACITY, "LDS image of %zu bytes too large (%d ice variants)", img.size(), nvar);
  HIPCHK(hipStreamSynchronize(c->stream));  // a launch in flight may still read the old image
  if (c->d_image) (void)hipFree(c->d_image), c->d_image = nullptr;
  HIPCHK(dalloc(&c->d_image, img.size() / 16));
  HIPCHK(hipMemcpy(c->d_image, img.data(), img.size(), hipMemcpyHostToDevice));
  c->lds_bytes = img.size();
  P.image_words = (int)(img.size() / 16);
  P.image = c->d_image;
  c->nvar = nvar;
  return PPC_HIP_OK;
}
}  // namespace

extern "C" {

const char* ppc_hip_last_error(void) { return g_err.c_str(); }

int ppc_hip_device_count(int32_t* count) {
  if (!count) return fail(PPC_HIP_E_ARG, "count is NULL");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  *count = n;
  return PPC_HIP_OK;
}

int ppc_hip_device_info(int32_t device, char* name, size_t name_len, int32_t* cus, uint64_t* mem) {
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (name && name_len > 0) {
    snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cus) *cus = prop.multiProcessorCount;
  if (mem) *mem = prop.totalGlobalMem;
  return PPC_HIP_OK;
}

int ppc_hip_open(int32_t device, const ppc_hip_config* cfg, ppc_hip_ctx** out) {
  if (!out) return fail(PPC_HIP_E_ARG, "ctx is NULL");
  *out = nullptr;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(PPC_HIP_E_ARG, "device %d not present (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));

  ppc_hip_ctx* c = new ppc_hip_ctx();
  c->device = device;
  c->cus = prop.multiProcessorCount;
  if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) c->wall_khz = 0;
  if (cfg) c->cfg = *cfg;
  ppc_hip_config& k = c->cfg;
  if (k.threads <= 0) k.threads = 256;
  if (k.threads % 64 != 0 || k.threads > 1024) {
    delete c;
    return fail(PPC_HIP_E_ARG, "threads per workgroup must be a multiple of 64 up to 1024 (got %d)", k.threads);
  }
  // 28 waves per CU (7 per SIMD: the occupancy the kernel's register budget is compiled for)
  if (k.blocks <= 0) k.blocks = c->cus * std::max(1, ppc::kWavesPerSimd * 4 * 64 / k.threads);
  if (k.refill <= 0) k.refill = 10;
  if (k.chunk > 1024) {
    delete c;
    return fail(PPC_HIP_E_ARG, "chunk %u above 1024", k.chunk);
  }
  if (k.refill > 64) k.refill = 64;
  if (k.max_bunches == 0) k.max_bunches = 1u << 20;
  if (k.max_hits == 0) k.max_hits = 1u << 24;
  c->lanes = k.blocks * k.threads;

  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) {
    if (r != hipSuccess && e == hipSuccess) e = r;
  };
  chk(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  chk(hipEventCreate(&c->ev0));
  chk(hipEventCreate(&c->ev1));
  chk(hipEventCreateWithFlags(&c->ev_copy, hipEventDisableTiming));
  chk(hipEventCreate(&c->ev_a0));
  chk(hipEventCreate(&c->ev_a1));
  chk(dalloc(&c->d_pz, k.max_bunches));
  chk(dalloc(&c->d_pstart, k.max_bunches + 1));
  chk(dalloc(&c->d_hits, k.max_hits));
  chk(dalloc(&c->d_hit_photon, k.max_hits));
  chk(dalloc(&c->d_ctr, 1));
  chk(dalloc(&c->d_rec, k.rec_max > 0 ? k.rec_max : 1));
  chk(dalloc(&c->d_rs, c->lanes));
  chk(dalloc(&c->d_rm, c->lanes));
  chk(hipHostMalloc((void**)&c->h_pz, sizeof(ppc_hip_bunch) * k.max_bunches, hipHostMallocDefault));
  chk(hipHostMalloc((void**)&c->h_pstart, sizeof(uint32_t) * (k.max_bunches + 1), hipHostMallocDefault));
  chk(hipHostMalloc((void**)&c->h_ctr, sizeof(ppc::Counters), hipHostMallocDefault));
  chk(hipHostMalloc((void**)&c->h_P, sizeof(KParams), hipHostMallocDefault));
  chk(dalloc(&c->d_P, 1));
  chk(dalloc(&c->d_bvar, k.max_bunches));
  chk(hipHostMalloc((void**)&c->h_bvar, sizeof(uint16_t) * k.max_bunches, hipHostMallocDefault));
  if (k.record_starts) chk(dalloc(&c->d_hphot, k.max_hits));
  chk(dalloc(&c->d_ctr_adj, 1));
  chk(hipHostMalloc((void**)&c->h_ctr_adj, sizeof(ppc::Counters), hipHostMallocDefault));
  if (e != hipSuccess) {
    int code = e == hipErrorOutOfMemory ? PPC_HIP_E_NOMEM : PPC_HIP_E_DEVICE;
    ppc_hip_close(c);
    return fail(code, "allocation failed: %s", hipGetErrorString(e));
  }
  e = scratch_address(device, &c->scratch);
  if (e != hipSuccess) {
    ppc_hip_close(c);
    return fail(PPC_HIP_E_DEVICE, "scratch area of device %d: %s", device, hipGetErrorString(e));
  }
  *out = c;
  return PPC_HIP_OK;
}

int ppc_hip_close(ppc_hip_ctx* c) {
  if (!c) return PPC_HIP_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : {(void*)c->d_pz, (void*)c->d_pstart, (void*)c->d_hits, (void*)c->d_hit_photon, (void*)c->d_ctr,
                  (void*)c->d_rec, (void*)c->d_rs, (void*)c->d_rm, (void*)c->d_image, (void*)c->d_lp, (void*)c->d_wv,
                  (void*)c->d_dom, (void*)c->d_domF, (void*)c->d_P, (void*)c->d_bvar, (void*)c->d_hphot, (void*)c->d_ctr_adj, (void*)c->d_clear, (void*)c->d_sensors,
                  (void*)c->d_pmt_dirs, (void*)c->d_rat, (void*)c->d_adoms, (void*)c->d_awvs, (void*)c->d_accepted,
                  (void*)c->d_adj_rec, (void*)c->d_adj_w, (void*)c->d_adj_dp, (void*)c->d_adj_grad})
    if (p) (void)hipFree(p);
  drop_graph(c);
  for (void* p : {(void*)c->h_pz, (void*)c->h_pstart, (void*)c->h_ctr, (void*)c->h_P, (void*)c->h_bvar, (void*)c->h_ctr_adj})
    if (p) (void)hipHostFree(p);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->ev_copy) (void)hipEventDestroy(c->ev_copy);
  if (c->ev_a0) (void)hipEventDestroy(c->ev_a0);
  if (c->ev_a1) (void)hipEventDestroy(c->ev_a1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PPC_HIP_OK;
}

int ppc_hip_lanes(const ppc_hip_ctx* c, int32_t* lanes) {
  if (!c || !lanes) return fail(PPC_HIP_E_ARG, "NULL argument");
  *lanes = c->lanes;
  return PPC_HIP_OK;
}

int ppc_hip_capacity(const ppc_hip_ctx* c, uint32_t* max_bunches, uint32_t* max_hits) {
  if (!c) return fail(PPC_HIP_E_ARG, "NULL context");
  if (max_bunches) *max_bunches = c->cfg.max_bunches;
  if (max_hits) *max_hits = c->cfg.max_hits;
  return PPC_HIP_OK;
}

int ppc_hip_clearance_map(const ppc_hip_tables* t, float step, float* map, size_t cap, int32_t* nx, int32_t* ny,
                          float* x0, float* y0) {
  if (!t || !nx || !ny || !x0 || !y0 || !(step > 0) || t->nsc < 1 || !t->sc)
    return fail(PPC_HIP_E_ARG, "clearance map needs strings and a positive step");
  float xl = t->sc[0].x, xh = xl, yl = t->sc[0].y, yh = yl;
  for (int i = 1; i < t->nsc; i++) {
    xl = std::min(xl, t->sc[i].x), xh = std::max(xh, t->sc[i].x);
    yl = std::min(yl, t->sc[i].y), yh = std::max(yh, t->sc[i].y);
  }
  *x0 = xl - 100.f, *y0 = yl - 100.f;
  *nx = (int)std::ceil((xh + 100.f - *x0) / step), *ny = (int)std::ceil((yh + 100.f - *y0) / step);
  if (!map) return PPC_HIP_OK;
  if (cap < (size_t)*nx * *ny) return fail(PPC_HIP_E_CAPACITY, "clearance map needs %d x %d entries", *nx, *ny);
  // a point of the square is at most half a diagonal from its centre; 0.05 m covers rounding.
  // Entries are capped at kClearCap (a smaller lower bound is still one), so each line only
  // visits the squares whose centres lie within kClearCap of its cylinder.
  const double slack = step * 0.70710678 + 0.05;
  std::vector<double> best((size_t)*nx * *ny, kClearCap);
  for (int k = 0; k < t->nsc; k++) {
    const ppc_hip_line& l = t->sc[k];
    const double rad = std::max<double>(l.r, t->hr), reach = kClearCap + rad;
    const int i0 = std::max(0, (int)std::floor((l.x - reach - *x0) / step));
    const int i1 = std::min(*nx - 1, (int)std::ceil((l.x + reach - *x0) / step));
    const int j0 = std::max(0, (int)std::floor((l.y - reach - *y0) / step));
    const int j1 = std::min(*ny - 1, (int)std::ceil((l.y + reach - *y0) / step));
    for (int j = j0; j <= j1; j++) {
      const double dy = *y0 + (j + 0.5) * step - l.y;
      for (int i = i0; i <= i1; i++) {
        const double dx = *x0 + (i + 0.5) * step - l.x;
        double& b = best[(size_t)j * *nx + i];
        b = std::min(b, std::sqrt(dx * dx + dy * dy) - rad);
      }
    }
  }
  for (size_t q = 0; q < best.size(); q++) map[q] = (float)(best[q] - slack);
  return PPC_HIP_OK;
}

int ppc_hip_upload(ppc_hip_ctx* c, const ppc_hip_tables* t) {
  if (!c || !t) return fail(PPC_HIP_E_ARG, "NULL argument");
  if (t->size < 1 || t->size > PPC_HIP_MAXLYS) return fail(PPC_HIP_E_ARG, "layers %d out of range", t->size);
  if (t->gsize < 0 || t->nsc < 0 || t->nsc > PPC_HIP_NSTR) return fail(PPC_HIP_E_ARG, "bad geometry sizes");
  if (t->rsize < c->lanes)
    return fail(PPC_HIP_E_CAPACITY, "not enough multipliers: %d for %d lanes", t->rsize, c->lanes);
  if (t->cn[0] < 0 || t->cn[0] > PPC_HIP_CX || t->cn[1] < 0 || t->cn[1] > PPC_HIP_CY)
    return fail(PPC_HIP_E_ARG, "cell grid %dx%d exceeds %dx%d", t->cn[0], t->cn[1], PPC_HIP_CX, PPC_HIP_CY);
  if (t->tmod != 0 && (t->lnum < 1 || t->lnum > PPC_HIP_LMAX || t->lpts < 0 || t->lpts + 1 > PPC_HIP_LYRS))
    return fail(PPC_HIP_E_ARG, "tilt table %d x %d out of range", t->lnum, t->lpts);
  if (t->tmod == 1 && t->lnum < 2) return fail(PPC_HIP_E_ARG, "1-d tilt needs at least 2 reference points");
  if (t->tmod == 2) {
    if (t->mnum[0] < 1 || t->mnum[0] > PPC_HIP_CTX || t->mnum[1] < 1 || t->mnum[1] > PPC_HIP_CTY)
      return fail(PPC_HIP_E_ARG, "2-d tilt grid %dx%d out of range", t->mnum[0], t->mnum[1]);
    for (int i = 0; i < t->mnum[0]; i++)
      for (int j = 0; j < t->mnum[1]; j++)
        if (t->mcol[i * PPC_HIP_CTY + j] >= t->lnum) return fail(PPC_HIP_E_ARG, "2-d tilt map index out of range");
  }
  if (t->tmod < 0 || t->tmod > 2) return fail(PPC_HIP_E_ARG, "tilt model %d unknown", t->tmod);
  // every cell list entry must name a line, every line a DOM range inside the geometry
  for (int i = 0; i < t->cn[0]; i++)
    for (int j = 0; j < t->cn[1]; j++)
      for (uint32_t k = t->is[i * PPC_HIP_CY + j], guard = 0; k != 0x8000u; guard++) {
        if (k >= PPC_HIP_NSTR || guard > PPC_HIP_NSTR) return fail(PPC_HIP_E_ARG, "corrupt cell list");
        uint32_t m = t->ls[k];
        if ((int)(m & 0x7fff) >= t->nsc) return fail(PPC_HIP_E_ARG, "cell list names line %u of %d", m & 0x7fff, t->nsc);
        k = m & 0x8000 ? 0x8000u : k + 1;
      }
  for (int i = 0; i < t->nsc; i++) {
    const ppc_hip_line& s = t->sc[i];
    if (s.n < 0 || s.max < 0 || s.n + s.max >= t->gsize || s.n > 0xffff || s.max > 0xffff)
      return fail(PPC_HIP_E_ARG, "line %d: DOMs %d..%d outside 0..%d", i, s.n, s.n + s.max, t->gsize - 1);
  }
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graph(c);  // a captured graph reads the tables it was captured with

  KParams& P = c->P;
  P = KParams{};
  P.size = t->size, P.dh = t->dh, P.rdh = t->rdh, P.hmin = t->hmin;
  P.ocv = t->ocv, P.sf = t->sf, P.g = t->g, P.gr = t->gr, P.xR = t->xR;
  P.SF = t->SF, P.G = t->G, P.GR = t->GR, P.hr = t->hr, P.hr2 = t->hr2, P.hs = t->hs, P.ha = t->ha;
  P.hifl = t->hifl;
  P.azx = t->azx, P.azy = t->azy, P.k1 = t->k1, P.k2 = t->k2, P.kz = t->kz, P.fr = t->fr;
  P.sum = t->sum;
  for (int i = 0; i < 12; i++) P.bfr[i] = t->bfr[i];
  P.rx = t->rx;
  P.cn0 = t->cn[0], P.cn1 = t->cn[1];
  P.cl0 = t->cl[0], P.cl1 = t->cl[1], P.crst0 = t->crst[0], P.crst1 = t->crst[1];
  P.cb00 = t->cb[0], P.cb01 = t->cb[1], P.cb10 = t->cb[2], P.cb11 = t->cb[3];
  P.tmod = t->tmod, P.vthk = t->vthk, P.lnum = t->lnum, P.lpts = t->lpts;
  P.lmin = t->lmin, P.lrdz = t->lrdz, P.r0 = t->r0, P.lnx = t->lnx, P.lny = t->lny;
  for (int i = 0; i < PPC_HIP_LMAX; i++) P.lr[i] = t->lr[i];
  P.mmin0 = t->mmin[0], P.mmin1 = t->mmin[1], P.mstp0 = t->mstp[0], P.mstp1 = t->mstp[1];
  P.mnum0 = t->mnum[0], P.mnum1 = t->mnum[1], P.mcut0 = t->mcut[0], P.mcut1 = t->mcut[1];
  P.nsc = t->nsc, P.gsize = t->gsize;
  {
    const float sfv[2] = {t->sf, t->SF}, gv[2] = {t->g, t->G};
    for (int k = 0; k < 2; k++) {
      P.hg_isf1[k] = 1.f / (1.f - sfv[k]);
      P.hg_a[k] = 1.f - gv[k] * gv[k];
      P.hg_b[k] = (1.f - gv[k]) * (1.f - gv[k]);
      P.hg_i2g[k] = 1.f / (2.f * gv[k]);
      P.hg_isf[k] = 1.f / sfv[k];
    }
  }

  // ---- LDS image ----
  std::vector<uint8_t> img;
  auto put = [&](const void* src, size_t bytes) {
    size_t off = (img.size() + 15) & ~size_t(15);
    img.resize(off + bytes);
    if (bytes) memcpy(img.data() + off, src, bytes);
    return (int)off;
  };
  P.scat_frame = 0;
  for (int i = 0; i < t->size; i++)
    if (t->az[i].k1 != 1.f || t->az[i].k2 != 1.f) P.scat_frame = 1;
  // the per-layer anisotropy table is read by the scattering frame, birefringence and the old
  // absorption anisotropy only: without them it stays out of the LDS image
  if (P.scat_frame || t->sum > 0 || t->hr > 0 || t->fr != 0) {
    // per layer (k1, k2, ra, rb), then (1/k1, 1/k2, 1/(k1 k2), k1 k2) for the scattering frame
    std::vector<float4> az(2 * t->size);
    for (int i = 0; i < t->size; i++) {
      const ppc_hip_aniz& a = t->az[i];
      const float okz = a.k1 * a.k2;
      az[i] = make_float4(a.k1, a.k2, a.ra, a.rb);
      az[t->size + i] = make_float4(1.f / a.k1, 1.f / a.k2, 1.f / okz, okz);
    }
    P.off_az = put(az.data(), sizeof(float4) * az.size());
  }
  {  // 1-d tilt reference distances, then the reciprocal of each interval (j-1, j)
    std::vector<float> lr(2 * PPC_HIP_LMAX, 0.f);
    for (int j = 0; j < PPC_HIP_LMAX; j++) lr[j] = t->lr[j];
    for (int j = 1; j < t->lnum && j < PPC_HIP_LMAX; j++)
      lr[PPC_HIP_LMAX + j] = 1.f / (t->lr[j] - t->lr[j - 1]);  // inf for a zero interval, as the division
    P.off_lr = put(lr.data(), sizeof(float) * lr.size());
  }
  P.off_mcol = put(t->mcol, sizeof(int32_t) * PPC_HIP_CTX * PPC_HIP_CTY);
  P.off_is = put(t->is, sizeof(uint16_t) * PPC_HIP_CX * PPC_HIP_CY);
  P.off_ls = put(t->ls, sizeof(uint16_t) * PPC_HIP_NSTR);
  {
    std::vector<float4> sc(2 * std::max(t->nsc, 1));
    for (int i = 0; i < t->nsc; i++) {
      const ppc_hip_line& s = t->sc[i];
      sc[2 * i] = make_float4(s.x, s.y, s.r, s.h);
      uint32_t bits = (uint32_t)s.n | ((uint32_t)s.max << 16);
      float fb;
      memcpy(&fb, &bits, 4);
      sc[2 * i + 1] = make_float4(s.d, s.dl, s.dh, fb);
    }
    P.off_sc = put(sc.data(), sizeof(float4) * sc.size());
  }
  // tilt table: rows k = 0..lpts are read (clamped), so lpts+1 entries per reference point
  const int lp_stride = t->tmod != 0 ? t->lpts + 1 : 1;
  const int lp_rows = t->tmod != 0 ? t->lnum : 0;
  std::vector<float2> lp(std::max(lp_rows * lp_stride, 1));
  for (int i = 0; i < lp_rows; i++)
    for (int k = 0; k < lp_stride; k++)
      lp[i * lp_stride + k] = make_float2(t->lp[(i * PPC_HIP_LYRS + k) * 2], t->lp[(i * PPC_HIP_LYRS + k) * 2 + 1]);
  P.lp_stride = lp_stride;
  {
    const size_t lp_bytes = lp.size() * sizeof(float2);
    // the 1-d tilt reads its table from global memory; the 2-d tilt's three columns per
    // step come from LDS when the table fits the budget of the tilt_lds setting
    const int mode = c->cfg.tilt_lds;
    P.lp_in_lds = t->tmod != 2 || mode < 0 ? 0 : lp_bytes <= (mode > 0 ? size_t(40960) : size_t(16384)) ? 1 : 0;
  }
  if (P.lp_in_lds) P.off_lp = put(lp.data(), sizeof(float2) * lp.size());
  img.resize((img.size() + 15) & ~size_t(15));
  c->img_base = img;  // the layer table (ice variants) is appended by set_layers

  for (void* p : {(void*)c->d_image, (void*)c->d_lp, (void*)c->d_wv, (void*)c->d_dom, (void*)c->d_domF})
    if (p) (void)hipFree(p);
  c->d_image = nullptr, c->d_lp = nullptr, c->d_wv = nullptr, c->d_dom = nullptr, c->d_domF = nullptr;
  {
    int rc = set_layers(c, t->lay, 1);
    if (rc != PPC_HIP_OK) return rc;
  }
  HIPCHK(dalloc(&c->d_lp, lp.size()));
  HIPCHK(hipMemcpy(c->d_lp, lp.data(), sizeof(float2) * lp.size(), hipMemcpyHostToDevice));
  HIPCHK(dalloc(&c->d_wv, PPC_HIP_WNUM));
  HIPCHK(hipMemcpy(c->d_wv, t->wv, sizeof(ppc_hip_wv) * PPC_HIP_WNUM, hipMemcpyHostToDevice));
  {
    std::vector<float4> dom(std::max(t->gsize, 1));
    std::vector<float> domF(std::max(t->gsize, 1), 1.f);
    bool uniform = true;
    for (int i = 0; i < t->gsize; i++) {
      const ppc_hip_dom& d = t->oms[i];
      dom[i] = make_float4(d.r[0], d.r[1], d.r[2], d.R);
      domF[i] = d.F;
      if (d.F != 1.f) uniform = false;
    }
    P.uniform_dom = uniform ? 1 : 0;
    HIPCHK(dalloc(&c->d_dom, dom.size()));
    HIPCHK(hipMemcpy(c->d_dom, dom.data(), sizeof(float4) * dom.size(), hipMemcpyHostToDevice));
    HIPCHK(dalloc(&c->d_domF, domF.size()));
    HIPCHK(hipMemcpy(c->d_domF, domF.data(), sizeof(float) * domF.size(), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(c->d_rm, t->rm, sizeof(uint32_t) * c->lanes, hipMemcpyHostToDevice));
  {  // clearance map of the string cylinders: squares of 1 m over the strings' footprint + 100 m
     // (of 2 m, 4 m, ... where a map of 1 m squares would exceed kClearMaxEntries)
    std::vector<float> cm;
    int nx = 0, ny = 0;
    float x0 = 0, y0 = 0;
    float step = 1.f;
    if (t->nsc > 0 && !c->cfg.no_clearance) {
      int need;
      while ((need = ppc_hip_clearance_map(t, step, nullptr, 0, &nx, &ny, &x0, &y0)) == PPC_HIP_OK &&
             (size_t)nx * ny > kClearMaxEntries)
        step *= 2;
      if (need < 0) return need;
      cm.resize((size_t)nx * ny);
      ppc_hip_clearance_map(t, step, cm.data(), cm.size(), &nx, &ny, &x0, &y0);
    }
    if (c->d_clear) (void)hipFree(c->d_clear), c->d_clear = nullptr;
    if (!cm.empty()) {
      HIPCHK(dalloc(&c->d_clear, cm.size()));
      HIPCHK(hipMemcpy(c->d_clear, cm.data(), sizeof(float) * cm.size(), hipMemcpyHostToDevice));
    }
    P.clear_map = c->d_clear;
    P.cm_x0 = x0, P.cm_y0 = y0, P.cm_rstep = 1.f / step, P.cm_nx = nx, P.cm_ny = ny;
  }

  P.lp = c->d_lp;
  P.wv = c->d_wv;
  P.dom = c->d_dom;
  P.domF = c->d_domF;
  P.rm = c->d_rm;
  P.rs = c->d_rs;
  P.pz = c->d_pz;
  P.pstart = c->d_pstart;
  P.hits = c->d_hits;
  P.hit_photon = c->d_hit_photon;
  P.hnum = c->cfg.max_hits;
  P.ctr = c->d_ctr;
  P.rec = c->d_rec;
  P.rec_stride = c->cfg.rec_stride;
  P.rec_max = c->cfg.rec_max;
  P.refill = c->cfg.refill;
  P.deterministic = c->cfg.deterministic;
  P.lanes = (uint32_t)c->lanes;
  P.hit_photons = c->d_hphot;
  P.hit_photons_cap = c->d_hphot ? c->cfg.max_hits : 0;
  c->uploaded = true;
  return ppc_hip_seed(c, 0);
}

int ppc_hip_seed(ppc_hip_ctx* c, uint32_t sv) {
  if (!c) return fail(PPC_HIP_E_ARG, "NULL context");
  std::vector<uint64_t> st(c->lanes, (uint64_t)(uint32_t)(362436069u + sv) | ((uint64_t)1234567u << 32));
  return ppc_hip_set_states(c, st.data(), c->lanes);
}

int ppc_hip_seed_photons(ppc_hip_ctx* c, uint64_t seed) {
  if (!c) return fail(PPC_HIP_E_ARG, "NULL context");
  c->photon_seed = seed;
  return PPC_HIP_OK;
}

int ppc_hip_set_states(ppc_hip_ctx* c, const uint64_t* states, int32_t lanes) {
  if (!c || !states || lanes != c->lanes) return fail(PPC_HIP_E_ARG, "states: need %d lanes", c ? c->lanes : 0);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->d_rs, states, sizeof(uint64_t) * lanes, hipMemcpyHostToDevice));
  return PPC_HIP_OK;
}

int ppc_hip_get_states(ppc_hip_ctx* c, uint64_t* states, int32_t lanes) {
  if (!c || !states || lanes != c->lanes) return fail(PPC_HIP_E_ARG, "states: need %d lanes", c ? c->lanes : 0);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(states, c->d_rs, sizeof(uint64_t) * lanes, hipMemcpyDeviceToHost));
  return PPC_HIP_OK;
}

namespace {

// check the bunches and copy them (and their ice variants) into pinned staging
int stage(ppc_hip_ctx* c, const ppc_hip_bunch* bunches, const uint16_t* variant, uint32_t nb, uint32_t num) {
  if (!c->uploaded) return fail(PPC_HIP_E_STATE, "launch before upload");
  if (nb > c->cfg.max_bunches) return fail(PPC_HIP_E_CAPACITY, "%u bunches > capacity %u", nb, c->cfg.max_bunches);
  if (num > 0 && (nb == 0 || !bunches)) return fail(PPC_HIP_E_ARG, "photons without bunches");
  // the photon counter runs past num by at most one chunk (<= 1024) per wave
  if ((uint64_t)num + (uint64_t)(c->lanes / 64 + 1) * 1024 > 0xffffffffull)
    return fail(PPC_HIP_E_CAPACITY, "%u photons in one launch is too many", num);
  HIPCHK(hipSetDevice(c->device));
  // the staging buffers may still feed the previous copy
  if (c->copy_pending) HIPCHK(hipEventSynchronize(c->ev_copy));
  uint64_t total = 0;
  for (uint32_t b = 0; b < nb; b++) {
    c->h_pstart[b] = (uint32_t)std::min<uint64_t>(total, 0xffffffffu);
    total += bunches[b].num;
  }
  if (total < num) return fail(PPC_HIP_E_ARG, "bunches hold %llu photons, launch asks for %u", (unsigned long long)total, num);
  for (uint32_t b = 0; b < nb; b++) {  // DOM indices the kernel dereferences
    const ppc_hip_bunch& p = bunches[b];
    if (p.type > 0 && p.u.l.ka > 999.f && p.n[3] != 0 && (p.u.l.fla < 0 || p.u.l.fla >= c->P.gsize))
      return fail(PPC_HIP_E_ARG, "bunch %u: flasher DOM %d outside the geometry", b, (int)p.u.l.fla);
    if (variant && variant[b] >= c->nvar)
      return fail(PPC_HIP_E_ARG, "bunch %u: ice variant %u of %d", b, (unsigned)variant[b], c->nvar);
  }
  c->h_pstart[nb] = (uint32_t)std::min<uint64_t>(total, 0xffffffffu);
  if (nb) memcpy(c->h_pz, bunches, sizeof(ppc_hip_bunch) * nb);
  if (nb && variant) memcpy(c->h_bvar, variant, sizeof(uint16_t) * nb);
  *c->h_P = c->P;
  c->h_P->nb = nb, c->h_P->num = num;
  c->h_P->photon_seed = c->photon_seed;
  c->h_P->bvar = variant ? c->d_bvar : nullptr;
  {  // claim chunks of photons per wave: large enough to keep the global counter cold, small
     // enough that the last chunks of a launch still spread over all waves: each wave
     // makes about 128 claims per launch, within 64..1024 photons per claim
    const uint32_t waves = (uint32_t)c->lanes / 64;
    uint64_t ch = num / ((uint64_t)waves * 128 + 1);
    c->h_P->chunk = c->cfg.chunk ? c->cfg.chunk : (uint32_t)std::min<uint64_t>(std::max<uint64_t>(ch, 64), 1024);
  }
  return PPC_HIP_OK;
}

// copies, counter reset and kernel on the context's stream (graph-capturable: no memset nodes)
int enqueue(ppc_hip_ctx* c, uint32_t nb, bool variants, bool timed) {
  if (nb) {
    HIPCHK(hipMemcpyAsync(c->d_pz, c->h_pz, sizeof(ppc_hip_bunch) * nb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_pstart, c->h_pstart, sizeof(uint32_t) * (nb + 1), hipMemcpyHostToDevice, c->stream));
    if (variants)
      HIPCHK(hipMemcpyAsync(c->d_bvar, c->h_bvar, sizeof(uint16_t) * nb, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(hipMemcpyAsync(c->d_P, c->h_P, sizeof(KParams), hipMemcpyHostToDevice, c->stream));
  if (timed) HIPCHK(hipEventRecord(c->ev_copy, c->stream));
  HIPCHK(ppc::launch_reset(c->d_ctr, c->stream));
  if (timed) HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (c->h_P->num > 0)
    HIPCHK(ppc::launch_propagate(*c->h_P, c->d_P, c->cfg.blocks, c->cfg.threads, c->lds_bytes, c->stream,
                                 c->cfg.count_events != 0, false));
  if (timed) HIPCHK(hipEventRecord(c->ev1, c->stream));
  return PPC_HIP_OK;
}


// device buffers of one adjoint call: replay records given by the caller, LLH weights,
// d(p)/d(a) then d(s1)/d(b), the FP64 gradient, the launch counters
struct AdjBuffers {
  ppc_hip_adj_record* rec;
  float* weight;
  float* dp;
  double* grad;
  ppc::Counters* ctr;
};

int adjoint_check(ppc_hip_ctx* c, const ppc_hip_bunch* bunches, uint32_t nb, uint32_t num,
                  const ppc_hip_adj_record* rec, uint32_t nrec, const float* weight, uint32_t nseg,
                  const float* dpda, const float* ds1db, double* grad) {
  if (!c || !grad || !dpda || !ds1db || (nrec && !rec) || (nseg && !weight)) return fail(PPC_HIP_E_ARG, "NULL argument");
  if (!c->uploaded) return fail(PPC_HIP_E_STATE, "adjoint before upload");
  if (c->P.hr > 0) return fail(PPC_HIP_E_ARG, "the adjoint does not model hole ice");
  if (!rec) {  // the hitting photons the last launch recorded
    if (!c->d_hphot) return fail(PPC_HIP_E_STATE, "context opened without record_starts");
    if (!c->fwd_waited) return fail(PPC_HIP_E_STATE, "wait for the forward launch before its adjoint");
    if (num != c->last_num) return fail(PPC_HIP_E_ARG, "adjoint of %u photons, the launch had %u", num, c->last_num);
  } else {
    for (uint32_t k = 0; k < nrec; k++)
      if (rec[k].lane >= (uint32_t)c->lanes || rec[k].photon >= num)
        return fail(PPC_HIP_E_ARG, "adjoint record %u: photon %u lane %u out of range", k, rec[k].photon,
                    rec[k].lane);
  }
  (void)bunches, (void)nb;
  return PPC_HIP_OK;
}

// the replay launch of an adjoint into the buffers B (checked by adjoint_check)
int adjoint_run(ppc_hip_ctx* c, const AdjBuffers& B, const ppc_hip_bunch* bunches, uint32_t nb, uint32_t num,
                const ppc_hip_adj_record* rec, uint32_t nrec, const float* weight, uint32_t nseg,
                const float* dpda, const float* ds1db, double* grad, ppc_hip_stats* st) {
  if (!rec) nrec = c->fwd_nphot;
  int rc = stage(c, bunches, nullptr, nb, num);
  if (rc != PPC_HIP_OK) return rc;
  const int size = c->P.size;
  // the stream is idle here (stage waited for the staging copy; synchronous calls below)
  HIPCHK(hipStreamSynchronize(c->stream));
  if (rec && nrec) HIPCHK(hipMemcpy(B.rec, rec, sizeof(ppc_hip_adj_record) * nrec, hipMemcpyHostToDevice));
  if (nseg) HIPCHK(hipMemcpy(B.weight, weight, sizeof(float) * nseg * c->P.gsize, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(B.dp, dpda, sizeof(float) * size, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(B.dp + size, ds1db, sizeof(float) * size, hipMemcpyHostToDevice));
  // the gradient starts at zero: cleared by a kernel (no memset node in a captured stream)
  HIPCHK(launch_zero(reinterpret_cast<uint32_t*>(B.grad), (uint32_t)(sizeof(double) * 2 * size / 4), c->stream));
  KParams& H = *c->h_P;
  H.adj_rec = rec ? B.rec : c->d_hphot, H.adj_nrec = nrec, H.adj_nseg = nseg;
  H.adj_weight = B.weight, H.adj_dpda = B.dp, H.adj_ds1db = B.dp + size;
  H.adj_grad = B.grad;
  H.ctr = B.ctr;            // the forward launch's counters and hits stay as they are
  H.hit_photons = nullptr;
  HIPCHK(hipMemcpyAsync(c->d_pz, c->h_pz, sizeof(ppc_hip_bunch) * nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_pstart, c->h_pstart, sizeof(uint32_t) * (nb + 1), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_P, c->h_P, sizeof(KParams), hipMemcpyHostToDevice, c->stream));
  HIPCHK(ppc::launch_reset(B.ctr, c->stream));
  HIPCHK(hipEventRecord(c->ev_a0, c->stream));
  if (nrec)
    HIPCHK(ppc::launch_propagate(H, c->d_P, c->cfg.blocks, c->cfg.threads, c->lds_bytes, c->stream, false, true));
  HIPCHK(hipEventRecord(c->ev_a1, c->stream));
  HIPCHK(hipMemcpyAsync(c->h_ctr_adj, B.ctr, sizeof(ppc::Counters), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(grad, B.grad, sizeof(double) * 2 * size, hipMemcpyDeviceToHost));
  if (nrec && c->h_ctr_adj->waves != (uint32_t)(c->lanes / 64))
    return fail(PPC_HIP_E_INCOMPLETE, "the adjoint launch ended in %u of its %u waves", c->h_ctr_adj->waves,
                (uint32_t)(c->lanes / 64));
  if (st) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev_a0, c->ev_a1));
    const ppc::Counters& k = *c->h_ctr_adj;
    *st = ppc_hip_stats{};
    st->photons = k.photons, st->steps = k.steps, st->killed = k.killed, st->nonfinite = k.nonfinite;
    st->device_ms = ms, st->lanes = c->lanes;
  }
  return PPC_HIP_OK;
}

}  // namespace

int ppc_hip_launch(ppc_hip_ctx* c, const ppc_hip_bunch* bunches, uint32_t nb, uint32_t num) {
  return ppc_hip_launch_batch(c, bunches, nullptr, nb, num);
}

int ppc_hip_launch_batch(ppc_hip_ctx* c, const ppc_hip_bunch* bunches, const uint16_t* variant, uint32_t nb,
                         uint32_t num) {
  if (!c) return fail(PPC_HIP_E_ARG, "NULL context");
  int rc = stage(c, bunches, variant, nb, num);
  if (rc != PPC_HIP_OK) return rc;
  rc = enqueue(c, nb, variant != nullptr, true);
  if (rc != PPC_HIP_OK) return rc;
  c->copy_pending = true;
  c->launched = true;
  c->fwd_waited = false;
  c->last_num = num;
  return PPC_HIP_OK;
}

int ppc_hip_set_layers(ppc_hip_ctx* c, const ppc_hip_layer* lay, int32_t nvar) {
  if (!c || !lay || nvar < 1 || nvar > 0xffff) return fail(PPC_HIP_E_ARG, "bad layer table");
  if (!c->uploaded) return fail(PPC_HIP_E_STATE, "set_layers before upload");
  HIPCHK(hipSetDevice(c->device));
  drop_graph(c);  // a captured graph reads the layer table it was captured with
  return set_layers(c, lay, nvar);
}

int ppc_hip_graph_capture(ppc_hip_ctx* c, const ppc_hip_bunch* bunches, const uint16_t* variant, uint32_t nb,
                          uint32_t num) {
  if (!c) return fail(PPC_HIP_E_ARG, "NULL context");
  int rc = stage(c, bunches, variant, nb, num);
  if (rc != PPC_HIP_OK) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->copy_pending = false;
  drop_graph(c);
  // the graph owns its bunches and parameter block, so that later launches (which
  // reuse the staging buffers) leave its replays unchanged; it holds kernels only
  const uint32_t nbs = std::max<uint32_t>(nb, 1);
  HIPCHK(dalloc(&c->g_pz, nbs));
  HIPCHK(dalloc(&c->g_pstart, nbs + 1));
  HIPCHK(dalloc(&c->g_bvar, nbs));
  HIPCHK(dalloc(&c->g_P, 1));
  KParams& H = *c->h_P;
  H.pz = c->g_pz, H.pstart = c->g_pstart;
  H.bvar = variant ? c->g_bvar : nullptr;
  if (nb) {
    HIPCHK(hipMemcpy(c->g_pz, c->h_pz, sizeof(ppc_hip_bunch) * nb, hipMemcpyHostToDevice));
    if (variant) HIPCHK(hipMemcpy(c->g_bvar, c->h_bvar, sizeof(uint16_t) * nb, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(c->g_pstart, c->h_pstart, sizeof(uint32_t) * (nb + 1), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->g_P, c->h_P, sizeof(KParams), hipMemcpyHostToDevice));
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = ppc::launch_reset(c->d_ctr, c->stream);
  if (e == hipSuccess && num > 0)
    e = ppc::launch_propagate(H, c->g_P, c->cfg.blocks, c->cfg.threads, c->lds_bytes, c->stream,
                              c->cfg.count_events != 0, false);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    HIPCHK(e);
    HIPCHK(e2);
  }
  c->graph = g;
  HIPCHK(hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
  c->graph_num = num;
  return PPC_HIP_OK;
}

int ppc_hip_graph_launch(ppc_hip_ctx* c) {
  if (!c) return fail(PPC_HIP_E_ARG, "NULL context");
  if (!c->graph_exec) return fail(PPC_HIP_E_STATE, "no captured graph");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  HIPCHK(hipGraphLaunch(c->graph_exec, c->stream));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  c->launched = true;
  c->fwd_waited = false;
  c->last_num = c->graph_num;
  return PPC_HIP_OK;
}

int ppc_hip_upload_accept(ppc_hip_ctx* c, const ppc_hip_accept_tables* t) {
  if (!c || !t) return fail(PPC_HIP_E_ARG, "NULL argument");
  if (t->nsensor < 0 || t->npmt < 0 || t->nacc < 0 || t->gsize < 0 || (t->nsensor && !t->sensors) ||
      (t->npmt && !t->pmt_dirs) || (t->nacc && !t->rat) || (t->gsize && !t->doms) || !t->wvs)
    return fail(PPC_HIP_E_ARG, "bad acceptance tables");
  for (int k = 0; k < t->nsensor; k++) {
    const ppc_hip_sensor& S = t->sensors[k];
    if (S.beta == -2 && !(S.Rz != 0 && S.Rr > 0))
      return fail(PPC_HIP_E_ARG, "sensor %d: wavelength-shifting tube of radius %g and half-length %g", k, S.Rr, S.Rz);
    if (S.def && (S.ns < 1 || S.ns > 16)) return fail(PPC_HIP_E_ARG, "sensor %d: %d coefficients", k, S.ns);
    if (!S.def && (S.npmt < 0 || S.pmt0 < 0 || S.pmt0 + S.npmt > t->npmt))
      return fail(PPC_HIP_E_ARG, "sensor %d: PMTs %d..%d of %d", k, S.pmt0, S.pmt0 + S.npmt, t->npmt);
  }
  for (int i = 0; i < t->gsize; i++) {
    const ppc_hip_dom_accept& d = t->doms[i];
    if (d.sensor >= t->nsensor || (d.sensor >= 0 && (d.acc < 0 || d.acc >= t->nacc)))
      return fail(PPC_HIP_E_ARG, "DOM %d: sensor %d / acceptance %d out of range", i, d.sensor, d.acc);
  }
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (void* p : {(void*)c->d_sensors, (void*)c->d_pmt_dirs, (void*)c->d_rat, (void*)c->d_adoms, (void*)c->d_awvs})
    if (p) (void)hipFree(p);
  c->d_sensors = nullptr, c->d_pmt_dirs = nullptr, c->d_rat = nullptr, c->d_adoms = nullptr, c->d_awvs = nullptr;
  c->accept_gsize = -1;
  HIPCHK(dalloc(&c->d_sensors, std::max(t->nsensor, 1)));
  HIPCHK(dalloc(&c->d_pmt_dirs, 3 * std::max(t->npmt, 1)));
  HIPCHK(dalloc(&c->d_rat, (size_t)PPC_HIP_WNUM * std::max(t->nacc, 1)));
  HIPCHK(dalloc(&c->d_adoms, std::max(t->gsize, 1)));
  HIPCHK(dalloc(&c->d_awvs, 2 * PPC_HIP_WNUM));
  if (t->nsensor) HIPCHK(hipMemcpy(c->d_sensors, t->sensors, sizeof(ppc_hip_sensor) * t->nsensor, hipMemcpyHostToDevice));
  if (t->npmt) HIPCHK(hipMemcpy(c->d_pmt_dirs, t->pmt_dirs, sizeof(float) * 3 * t->npmt, hipMemcpyHostToDevice));
  if (t->nacc) HIPCHK(hipMemcpy(c->d_rat, t->rat, sizeof(float) * PPC_HIP_WNUM * t->nacc, hipMemcpyHostToDevice));
  if (t->gsize) HIPCHK(hipMemcpy(c->d_adoms, t->doms, sizeof(ppc_hip_dom_accept) * t->gsize, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_awvs, t->wvs, sizeof(float) * 2 * PPC_HIP_WNUM, hipMemcpyHostToDevice));
  c->accept_gsize = t->gsize;
  return PPC_HIP_OK;
}

int ppc_hip_accept(ppc_hip_ctx* c, uint64_t seed, ppc_hip_accepted* out, uint32_t cap, uint32_t* n) {
  if (!c || !n) return fail(PPC_HIP_E_ARG, "NULL argument");
  if (c->accept_gsize < 0) return fail(PPC_HIP_E_STATE, "no acceptance tables uploaded");
  if (!c->launched) return fail(PPC_HIP_E_STATE, "acceptance without launch");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(c->h_ctr, c->d_ctr, sizeof(ppc::Counters), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const uint32_t stored = std::min<uint32_t>(c->h_ctr->hidx, c->cfg.max_hits);
  const uint32_t m = std::min(stored, cap);
  if (m > 0) {
    if (!c->d_accepted) HIPCHK(dalloc(&c->d_accepted, c->cfg.max_hits));
    ppc::AcceptParams A{c->d_sensors, c->d_pmt_dirs, c->d_rat, c->d_adoms, c->d_awvs, c->accept_gsize};
    HIPCHK(ppc::launch_accept(A, c->d_hits, c->d_hit_photon, m, seed, c->d_accepted, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (out) HIPCHK(hipMemcpy(out, c->d_accepted, sizeof(ppc_hip_accepted) * m, hipMemcpyDeviceToHost));
  }
  *n = m;
  return PPC_HIP_OK;
}

int ppc_hip_hit_photons(ppc_hip_ctx* c, ppc_hip_adj_record* rec, uint32_t cap, uint32_t* n) {
  if (!c || !n) return fail(PPC_HIP_E_ARG, "NULL argument");
  if (!c->d_hphot) return fail(PPC_HIP_E_STATE, "context opened without record_starts");
  if (!c->launched) return fail(PPC_HIP_E_STATE, "hit photons without launch");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(c->h_ctr, c->d_ctr, sizeof(ppc::Counters), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  uint32_t m = std::min<uint32_t>(std::min<uint32_t>(c->h_ctr->nphot, c->cfg.max_hits), cap);
  if (m > 0 && rec) HIPCHK(hipMemcpy(rec, c->d_hphot, sizeof(ppc_hip_adj_record) * m, hipMemcpyDeviceToHost));
  *n = m;
  return PPC_HIP_OK;
}

int ppc_hip_adjoint(ppc_hip_ctx* c, const ppc_hip_bunch* bunches, uint32_t nb, uint32_t num,
                    const ppc_hip_adj_record* rec, uint32_t nrec, const float* weight, uint32_t nseg,
                    const float* dpda, const float* ds1db, double* grad, ppc_hip_stats* st) {
  int rc = adjoint_check(c, bunches, nb, num, rec, nrec, weight, nseg, dpda, ds1db, grad);
  if (rc != PPC_HIP_OK) return rc;
  HIPCHK(hipSetDevice(c->device));
  // the adjoint's buffers grow on demand and stay with the context
  if (rec && nrec > c->adj_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->d_adj_rec) (void)hipFree(c->d_adj_rec);
    c->d_adj_rec = nullptr;
    HIPCHK(dalloc(&c->d_adj_rec, nrec));
    c->adj_cap = nrec;
  }
  const size_t wn = (size_t)std::max<uint32_t>(nseg, 1) * c->P.gsize;
  if (wn > c->adj_wcap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->d_adj_w) (void)hipFree(c->d_adj_w);
    c->d_adj_w = nullptr;
    HIPCHK(dalloc(&c->d_adj_w, wn));
    c->adj_wcap = wn;
  }
  if (!c->d_adj_dp) HIPCHK(dalloc(&c->d_adj_dp, 2 * PPC_HIP_MAXLYS));
  if (!c->d_adj_grad) HIPCHK(dalloc(&c->d_adj_grad, 2 * PPC_HIP_MAXLYS));
  AdjBuffers B{c->d_adj_rec, c->d_adj_w, c->d_adj_dp, c->d_adj_grad, c->d_ctr_adj};
  return adjoint_run(c, B, bunches, nb, num, rec, nrec, weight, nseg, dpda, ds1db, grad, st);
}

int ppc_hip_adjoint_noalloc(ppc_hip_ctx* c, const ppc_hip_bunch* bunches, uint32_t nb, uint32_t num,
                            const ppc_hip_adj_record* rec, uint32_t nrec, const float* weight, uint32_t nseg,
                            const float* dpda, const float* ds1db, double* grad, ppc_hip_stats* st) {
  int rc = adjoint_check(c, bunches, nb, num, rec, nrec, weight, nseg, dpda, ds1db, grad);
  if (rc != PPC_HIP_OK) return rc;
  if (!c->scratch) return fail(PPC_HIP_E_STATE, "no scratch area on device %d", c->device);
  if ((uint64_t)std::max<uint32_t>(nseg, 1) * c->P.gsize > kScratchWeights)
    return fail(PPC_HIP_E_CAPACITY, "%u segments x %d DOMs of weights exceed the scratch area's %u", nseg,
                c->P.gsize, kScratchWeights);
  if (rec && nrec > kScratchRecords)
    return fail(PPC_HIP_E_CAPACITY, "%u replay records exceed the scratch area's %u", nrec, kScratchRecords);
  HIPCHK(hipSetDevice(c->device));
  AdjScratch* S = c->scratch;
  AdjBuffers B{S->rec, S->weight, S->dp, S->grad, &S->ctr};
  return adjoint_run(c, B, bunches, nb, num, rec, nrec, weight, nseg, dpda, ds1db, grad, st);
}

int ppc_hip_wait(ppc_hip_ctx* c, ppc_hip_stats* st) {
  if (!c) return fail(PPC_HIP_E_ARG, "NULL context");
  if (!c->launched) return fail(PPC_HIP_E_STATE, "wait without launch");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(c->h_ctr, c->d_ctr, sizeof(ppc::Counters), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  {  // every wave ran to its end and every photon was started (cf. the MP check of gpu/ppc.cu:266);
     // a launch of no photons runs no kernel
    const ppc::Counters& k = *c->h_ctr;
    const uint32_t waves = c->last_num > 0 ? (uint32_t)(c->lanes / 64) : 0u;
    if (k.waves != waves) return fail(PPC_HIP_E_INCOMPLETE, "the launch ended in %u of its %u waves", k.waves, waves);
    if (k.photons != c->last_num)
      return fail(PPC_HIP_E_INCOMPLETE, "the launch started %llu of its %u photons", k.photons, c->last_num);
  }
  if (st) {
    const ppc::Counters& k = *c->h_ctr;
    st->photons = k.photons;
    st->steps = k.steps;
    st->hits = k.hidx;
    st->stored = std::min<uint64_t>(k.hidx, c->cfg.max_hits);
    st->killed = k.killed;
    st->nonfinite = k.nonfinite;
    st->device_ms = ms;
    st->lanes = c->lanes;
    st->crossings = k.crossings, st->strings = k.strings, st->doms = k.doms, st->scatters = k.scatters;
    st->hit_photons = k.nphot;
    st->span_min_ms = st->span_max_ms = 0.f;
    if (k.waves > 0 && c->wall_khz > 0) {
      const unsigned long long t0 = ~k.t_start_c, e0 = ~k.t_end_min_c, e1 = k.t_end_max;
      st->span_min_ms = (float)((double)(e0 - t0) / c->wall_khz);
      st->span_max_ms = (float)((double)(e1 - t0) / c->wall_khz);
    }
  }
  c->fwd_nphot = c->d_hphot ? std::min<uint32_t>(c->h_ctr->nphot, c->cfg.max_hits) : 0;
  c->fwd_waited = true;
  return PPC_HIP_OK;
}

int ppc_hip_hits(ppc_hip_ctx* c, ppc_hip_hit* hits, uint32_t* photon, uint32_t cap, uint32_t* n) {
  if (!c || !n) return fail(PPC_HIP_E_ARG, "NULL argument");
  if (!c->launched) return fail(PPC_HIP_E_STATE, "hits without launch");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(c->h_ctr, c->d_ctr, sizeof(ppc::Counters), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  uint32_t stored = std::min<uint32_t>(c->h_ctr->hidx, c->cfg.max_hits);
  uint32_t m = std::min(stored, cap);
  if (m > 0 && hits) HIPCHK(hipMemcpy(hits, c->d_hits, sizeof(ppc_hip_hit) * m, hipMemcpyDeviceToHost));
  if (m > 0 && photon) HIPCHK(hipMemcpy(photon, c->d_hit_photon, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  *n = m;
  return PPC_HIP_OK;
}

int ppc_hip_records(ppc_hip_ctx* c, ppc_hip_record* rec, uint32_t cap, uint32_t* n) {
  if (!c || !n) return fail(PPC_HIP_E_ARG, "NULL argument");
  if (!c->launched) return fail(PPC_HIP_E_STATE, "records without launch");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  uint32_t total = 0;
  if (c->cfg.rec_stride > 0) total = std::min<uint32_t>((c->last_num + c->cfg.rec_stride - 1) / c->cfg.rec_stride, c->cfg.rec_max);
  uint32_t m = std::min(total, cap);
  if (m > 0 && rec) HIPCHK(hipMemcpy(rec, c->d_rec, sizeof(ppc_hip_record) * m, hipMemcpyDeviceToHost));
  *n = m;
  return PPC_HIP_OK;
}

}  // extern "C"
