// frontend.h -- internal interface between the xppc front-end and the command line.
#ifndef XPPC_FRONTEND_H
#define XPPC_FRONTEND_H

#include "detector.h"

namespace xppc {

// output mode of print(): HIT lines on stdout (command line) or the hitz map (library)
extern bool cli_mode;

void f2k();                                                           // gpu/f2k.cxx:555-583
void flasher(int str, int dom, unsigned long long num, int itr);     // gpu/f2k.cxx:768-783
float host_zshift(const ppc_hip_tables& d, const float r[4], float& dz);  // gpu/pro.cu:168-264 on the host
void list_devices();

// bunch templates as the front-end builds them (used by the C facade)
ppc_hip_bunch flasher_bunch(int str, int dom);
// the bunches addp_clst() and addp_mopo() build for direction n (num left as set)
ppc_hip_bunch clst_bunch(float rx, float ry, float rz, float t, float nx, float ny, float nz, float dr, float beta);
double mopo_bunch(float rx, float ry, float rz, float t, float nx, float ny, float nz, float light, float length,
                  float beta, float tau, float fr, ppc_hip_bunch& out);
double particle_bunch(float rx, float ry, float rz, float t, float E, float dr, int type, float nx, float ny,
                      float nz, ppc_hip_bunch& out);

struct BackendInfo {
  int lanes, threads;
  unsigned pmax, pmxo, hnum;
  double device_ms;
  unsigned long long photons, hits, steps;
};
BackendInfo backend_info();
// the tables of print()'s acceptance for the device (ppc_hip_upload_accept)
const ppc_hip_accept_tables* accept_tables();
// print()'s stable sort of accepted photons by hitz key (multi-threaded)
void sort_by_key(std::vector<std::pair<ihit, pout>>& b);

}  // namespace xppc

#endif
