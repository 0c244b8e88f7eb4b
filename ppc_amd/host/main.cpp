// main.cpp -- the ppc command line (same argv modes and output as gpu/ppc.cu:514-599).
//
//   ppc                       list devices and usage
//   ppc -  [x y]              ice table at the middle wavelength bin, tilt-corrected depths
//   ppc =                     zshift for "x y z" points read from stdin
//   ppc _                     tilt surfaces on a 3 m x 3 m x 6 m grid
//   ppc [device]              f2k particles from stdin
//   ppc str om num[*itr] dev  flasher
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "frontend.h"

using namespace xppc;

static float zdh;

static float zshift_at(float x, float y, float z) {
  float r[4] = {x, y, z, 0};
  zdh = det.t.dh;
  return host_zshift(det.t, r, zdh);
}

int main(int argc, char* argv[]) {
  cli_mode = true;
  start();
  if (argc <= 1) {
    list_devices();
    fprintf(stderr,
            "Use: %s [device] (f2k muons)\n"
            "     %s [str] [om] [num] [device] (flasher)\n",
            argv[0], argv[0]);
  } else if (0 == strcmp(argv[1], "-")) {
    initialize(1.f);
    const ppc_hip_wv& w = det.wv[WNUM / 2];
    std::cerr << "For wavelength=" << det.wvs[w.wvl].w << " [nm]  np=" << (1 / w.coschr) << "  cm=" << 1 / w.ocm
              << " [m/ns]" << std::endl;
    float x = 0, y = 0;
    if (argc == 4) x = atof(argv[2]), y = atof(argv[3]);
    for (int i = 0; i < det.t.size; i++) {
      float z = det.t.hmin + det.t.dh * i, rz = z;
      for (int j = 0; j < 10; j++) rz = z + zshift_at(x, y, rz);
      z = rz;
      std::cout << z << " " << det.abs_tab[(WNUM / 2) * MAXLYS + i] << " "
                << det.sca_tab[(WNUM / 2) * MAXLYS + i] * (1 - det.t.g) << " " << det.az[i].ra * det.t.sum << std::endl;
    }
  } else if (0 == strcmp(argv[1], "=")) {
    initialize(1.f);
    const ppc_hip_wv& w = det.wv[WNUM / 2];
    std::cerr << "For wavelength=" << det.wvs[w.wvl].w << " [nm]  np=" << (1 / w.coschr) << "  cm=" << 1 / w.ocm
              << " [m/ns]" << std::endl;
    std::string in;
    while (getline(std::cin, in)) {
      float x, y, z;
      if (3 == sscanf(in.c_str(), "%f %f %f", &x, &y, &z)) {
        float dz = zshift_at(x, y, z);
        std::cout << in << " " << dz << " " << zdh << std::endl;
      }
    }
  } else if (0 == strcmp(argv[1], "_")) {
    initialize(1.f);
    for (float x = -750.f; x < 751.f; x += 3.f)
      for (float y = -750.f; y < 751.f; y += 3.f)
        for (float z = -750.f; z < 751.f; z += 6.f) {
          float rz = z;
          for (int j = 0; j < 10; j++) rz = z + zshift_at(x, y, rz);
          std::cout << z << " " << x << " " << y << " " << (rz - z) << "\n";
        }
  } else if (argc <= 2) {
    int device = atoi(argv[1]);
    initialize(1.f);
    choose(device);
    fprintf(stderr, "Processing f2k muons from stdin on device %d\n", device);
    f2k();
  } else {
    int str = 0, dom = 0, device = 0, itr = 0;
    unsigned long long num = 1000000ULL;
    if (argc > 1) str = atoi(argv[1]);
    if (argc > 2) dom = atoi(argv[2]);
    if (argc > 3) {
      num = (unsigned long long)atof(argv[3]);
      const char* sub = strchr(argv[3], '*');
      if (sub != NULL) itr = (int)atof(++sub);
    }
    if (argc > 4) device = atoi(argv[4]);
    initialize(1.f);
    choose(device);
    fprintf(stderr, "Running flasher simulation on device %d\n", device);
    flasher(str, dom, num, itr);
  }
  stop();
}
