"""The drop-in boundary, checked without a GPU.

* libppc_hip.so and libxppc.so load and export every function that
  include/ppc_hip.h and include/xppc_c.h declare;
* a C++ program written against the xppc interface (include/ppc.h, and the
  reference's own llh/ppc.h where the reference tree is present) compiles,
  links against libxppc.so and starts: every function and global the
  reference's consumers (llh, the IceTray module) use resolves.
"""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "ppc_amd", "lib")
REF_PPC_H = "/root/reference/llh/ppc.h"


def declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M)) - {"typedef"})


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIB, lib)], capture_output=True, text=True,
                         check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


@pytest.mark.parametrize("header,lib", [("ppc_hip.h", "libppc_hip.so"), ("xppc_c.h", "libxppc.so")])
def test_c_abi_exports(header, lib):
    names = declared(header)
    assert len(names) >= 8
    syms = exported(lib)
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    h = ctypes.CDLL(os.path.join(LIB, lib))
    for n in names:
        getattr(h, n)


def test_abi_struct_sizes():
    """the boundary structs keep the reference's sizes: bunch 64 B, hit 32 B"""
    from ppc_amd import api

    assert ctypes.sizeof(api.Bunch) == 64
    assert ctypes.sizeof(api.Hit) == 32


_SIZES = r'''
#include <stdio.h>
#include <stddef.h>
#include "ppc_hip.h"
int main(void) {
  printf("bunch %zu\nhit %zu\nconfig %zu\nstats %zu\nrecord %zu\nadj_record %zu\nlayer %zu\naccepted %zu\n"
         "sensor %zu\ndom_accept %zu\naccepted_dt %zu\n",
         sizeof(ppc_hip_bunch), sizeof(ppc_hip_hit), sizeof(ppc_hip_config), sizeof(ppc_hip_stats),
         sizeof(ppc_hip_record), sizeof(ppc_hip_adj_record), sizeof(ppc_hip_layer), sizeof(ppc_hip_accepted),
         sizeof(ppc_hip_sensor), sizeof(ppc_hip_dom_accept), offsetof(ppc_hip_accepted, dt));
  return 0;
}
'''


def test_python_mirrors_match_the_header():
    """the ctypes structures and numpy record types of the Python binding have the sizes of
    the C structures they mirror (a C program built against include/ppc_hip.h prints them)"""
    from ppc_amd import api

    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "sizes.c"), os.path.join(d, "sizes")
        open(src, "w").write(_SIZES)
        subprocess.run(["gcc", "-std=c99", "-I", INC, src, "-o", exe], check=True, capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    size = {l.split()[0]: int(l.split()[1]) for l in out.splitlines()}
    assert size["bunch"] == ctypes.sizeof(api.Bunch) == 64
    assert size["hit"] == ctypes.sizeof(api.Hit) == api.HIT_DTYPE.itemsize == 32
    assert size["config"] == ctypes.sizeof(api.Config)
    assert size["stats"] == ctypes.sizeof(api.Stats)
    assert size["record"] == api.RECORD_DTYPE.itemsize
    assert size["adj_record"] == api.ADJ_RECORD_DTYPE.itemsize
    assert size["layer"] == api.LAYER_DTYPE.itemsize
    assert size["accepted"] == api.ACCEPTED_DTYPE.itemsize
    assert size["accepted_dt"] == api.ACCEPTED_DTYPE.fields["dt"][1]


CONSUMER = r'''
#include <cstddef>
#include <cmath>
#include <map>
#include <string>
#include <utility>
#include <vector>
#include "ppc.h"

int main(int argc, char**) {
  if (argc > 5) {  // referenced for the linker only
    xppc::start(); xppc::stop(); xppc::choose(0); xppc::ini(); xppc::fin(); xppc::eout();
    xppc::flone(1); xppc::getMaxBunchSize(); xppc::getWorkgroupSize(); xppc::getTotalDeviceTime();
    xppc::initialize(1.f); xppc::flset(1, 1); xppc::fldir();
    float m[3] = {0, 0, 1}, n[3] = {1, 0, 0};
    xppc::flshift(m, n);
    xppc::set_res(1.f); xppc::set_res(1.f, 1.f); xppc::efin();
    xppc::sett(0, 0, 1, std::make_pair(0, 0ULL), 0);
    xppc::addp(0, 0, 0, 0, 1, 0, 0);
    xppc::addp_clst(0, 0, 0, 0, 1, 0, 1);
    xppc::addp_mopo(0, 0, 0, 0, 1, 1, 1, 0, 1);
    float r = xppc::xrnd() + xppc::grnd();
    return (int)(xppc::i3oms.size() + xppc::omts.size() + xppc::hvs.size() + xppc::rdes.size() +
                 xppc::hitz.size() + xppc::nextgen + r);
  }
  return 0;
}
'''


@pytest.mark.parametrize("header_dir", [INC, os.path.dirname(REF_PPC_H)])
def test_xppc_consumer_links(header_dir):
    if not os.path.exists(os.path.join(header_dir, "ppc.h")):
        pytest.skip("reference header not present")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "consumer.cpp")
        open(src, "w").write(CONSUMER)
        exe = os.path.join(d, "consumer")
        r = subprocess.run(["g++", "-std=c++11", "-I", header_dir, src, "-o", exe, "-L", LIB, "-lxppc",
                            "-Wl,-rpath," + LIB], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        assert subprocess.run([exe], timeout=60).returncode == 0


def test_cli_usage_without_device():
    from ppc_amd import _native

    r = subprocess.run([_native.ppc_exe()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "flasher" in r.stderr


def test_c_abi_rejects_bad_arguments_without_a_device():
    """argument checks come before any device work, and there is no CPU fallback:
    without a GPU opening a context fails with a status and a message"""
    from ppc_amd import api

    lib = api.hip()
    c_p = ctypes.c_void_p
    n = ctypes.c_uint32()
    assert lib.ppc_hip_open(0, None, None) < 0
    assert lib.ppc_hip_upload(None, None) < 0
    assert lib.ppc_hip_launch(None, None, 0, 0) < 0
    assert lib.ppc_hip_wait(None, None) < 0
    assert lib.ppc_hip_hits(None, None, None, 0, ctypes.byref(n)) < 0
    assert lib.ppc_hip_set_layers(None, None, 1) < 0
    assert lib.ppc_hip_graph_launch(None) < 0
    assert lib.ppc_hip_accept(None, 0, None, 0, ctypes.byref(n)) < 0
    assert lib.ppc_hip_upload_accept(None, None) < 0
    assert lib.ppc_hip_adjoint(None, None, 0, 0, None, 0, None, 0, None, None, None, None) < 0
    assert lib.ppc_hip_last_error()
    import torch

    if not torch.cuda.is_available():
        ctx = c_p()
        cfg = api.Config()
        assert lib.ppc_hip_open(0, ctypes.byref(cfg), ctypes.byref(ctx)) < 0
        assert b"" != lib.ppc_hip_last_error()
