"""The product's table builder (ppc_amd/host/ice.cpp) against the reference.

* field by field against the tables the reference's own ini::set() builds
  (oracle/_ref, where the reference tree is present);
* the `ppc -` (also at a position), `ppc =` and `ppc _` outputs of the product
  command line against the reference's outputs committed in tests/golden/;
* the MWC multiplier search against the reference's rnd.txt.
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_util as U

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
ROOT = os.path.dirname(HERE)
REFICE = "/root/reference/ice"

ENVS = {"c1": {"WFLA": "405", "FWID": "10"}, "c2": {}, "c2x": {}, "c1ng": {"WFLA": "405", "FWID": "10"},
        "c1wom": {"WFLA": "405", "FWID": "10"}, "c2xb": {}}


def _child(code, *args, env=None):
    """run in a fresh process: the table builders keep process-wide state"""
    e = dict(os.environ)
    for k in ("WFLA", "FWID", "VTHK"):
        e.pop(k, None)
    e.update(env or {})
    e["DTMN"] = "1"
    e["PYTHONPATH"] = ROOT + os.pathsep + HERE
    r = subprocess.run([sys.executable, "-c", code] + list(args), env=e, capture_output=True, text=True, timeout=300,
                       cwd=HERE)
    assert r.returncode == 0, r.stderr[-3000:]
    return r.stdout


_CMP = r'''
import sys, numpy as np
sys.path.insert(0, sys.argv[2])
import oracle_util as U
from ppc_amd import api
R = U.ref_tables(sys.argv[1])
D = api.Detector(sys.argv[1])
P = U.product_tables(D)
skip = {"rsize"}
for k in U.TABLE_FIELDS:
    if k in skip: continue
    a, b = bytes(R.field(k)), bytes(P.field(k))
    if a == b: continue
    if k in ("sca", "abs_"):
        fa, fb = np.frombuffer(a, np.float32), np.frombuffer(b, np.float32)
        rel = np.max(np.abs(fa - fb) / np.maximum(np.abs(fa), 1e-30))
        print("%s-maxrel" % k, rel)
        continue
    if k == "wvs_f":
        fa, fb = np.frombuffer(a, np.float32), np.frombuffer(b, np.float32)
        if (fa[:-1] == fb[:-1]).all(): continue
    print("DIFF", k)
'''


@pytest.mark.ref
@pytest.mark.parametrize("cfg", ["c1", "c2", "c2x", "c1ng", "c1wom", "c2xb"])
def test_tables_match_reference(cfg):
    if not U.ref_available():
        pytest.skip("oracle/_ref not built (no reference tree)")
    from ppc_amd import configs

    out = _child(_CMP, configs.ensure(cfg), HERE, env=ENVS[cfg])
    assert "DIFF" not in out, out
    # the dense scattering and absorption tables are computed in float with the same
    # formulas, both built with fast-math: the compilers may contract and reassociate
    # them differently (1 ulp in the synthetic ices; the absorption table's dust term,
    # zero by default, is nonzero under ABSM=1)
    for line in out.splitlines():
        if line.split()[0].endswith("-maxrel"):
            assert float(line.split()[1]) < 1e-6


# table-level switches of the environment (gpu/ini.cxx:629-729, 829-845, 1359-1363, 1608-1640)
ENV_SWITCHES = [{"BFRM": "1"}, {"BFRM": "2"}, {"BFRA": "0.8", "BFRB": "1.3"}, {"ABSM": "0"}, {"ABSM": "1"},
                {"GECO": "1 30 0.5 -0.3"}]


@pytest.mark.ref
@pytest.mark.parametrize("env", ENV_SWITCHES, ids=lambda e: "-".join("%s=%s" % kv for kv in e.items()))
@pytest.mark.parametrize("ice", ["c2x", "spice_bfr-v2"])
def test_tables_match_reference_under_switches(ice, env):
    """the same field-by-field comparison with the birefringence, absorption-table and
    DOM-position switches of the environment set"""
    if not U.ref_available():
        pytest.skip("oracle/_ref not built (no reference tree)")
    from ppc_amd import configs

    path = configs.ensure(ice) if ice.startswith("c") else os.path.join(REFICE, ice)
    if not os.path.isdir(path):
        pytest.skip("reference ice directories not present")
    out = _child(_CMP, path, HERE, env=env)
    assert "DIFF" not in out, out
    for line in out.splitlines():
        if line.split()[0].endswith("-maxrel"):
            assert float(line.split()[1]) < 1e-6


def _overrides(case, tmp):
    """(tables directory, environment) of a directory-override case (gpu/ini.cxx:595-625,
    1355-1363): the ice model, tilt model, nextgen sensors, hole ice or spectrum file taken
    from somewhere other than the tables directory"""
    from ppc_amd import configs

    if case == "ICEMODELDIR":  # the 86-string ice's layers and tilt under the 7-string geometry
        return configs.ensure("c1"), {"ICEMODELDIR": configs.ensure("c2"), "WFLA": "405"}
    if case == "TILTMODELDIR":  # the 2-d tilt map with the 1-d tilt ice's layers
        return configs.ensure("c2"), {"TILTMODELDIR": configs.ensure("c2x")}
    if case == "NEXTGENDIR":
        return configs.ensure("c1"), {"NEXTGENDIR": configs.ensure("c1ng"), "WFLA": "405"}
    if case == "PPCHOLEICE":
        return configs.ensure("c1"), {"PPCHOLEICE": os.path.join(configs.ensure("c2x"), "as.dat"), "WFLA": "405"}
    if case == "FLWL":  # another emission spectrum file, wv.<FLWL>, in the tables directory
        import shutil

        d = os.path.join(tmp, "c2flwl")
        shutil.copytree(configs.ensure("c2"), d)
        w = np.loadtxt(os.path.join(d, "wv.dat"))
        w[:, 1] = w[:, 1] ** 2 / w[-1, 1]  # a steeper cumulative spectrum
        np.savetxt(os.path.join(d, "wv.alt"), w, fmt="%.6f")
        return d, {"FLWL": "alt"}
    raise KeyError(case)


@pytest.mark.ref
@pytest.mark.parametrize("case", ["ICEMODELDIR", "TILTMODELDIR", "NEXTGENDIR", "PPCHOLEICE", "FLWL"])
def test_tables_match_reference_under_directory_overrides(case, tmp_path):
    """the field-by-field comparison with parts of the configuration read from elsewhere"""
    if not U.ref_available():
        pytest.skip("oracle/_ref not built (no reference tree)")
    path, env = _overrides(case, str(tmp_path))
    out = _child(_CMP, path, HERE, env=env)
    assert "DIFF" not in out, out
    for line in out.splitlines():
        if line.split()[0].endswith("-maxrel"):
            assert float(line.split()[1]) < 1e-6
    # the override took effect: the product's tables differ from those of the plain directory
    plain = dict(env)
    plain.pop(case)
    assert _child(_HASH, path, env=env) != _child(_HASH, path, env=plain)


_HASH = r'''
import sys, hashlib
import oracle_util as U
from ppc_amd import api
P = U.product_tables(api.Detector(sys.argv[1]))
h = hashlib.sha256()
for k in U.TABLE_FIELDS:
    h.update(bytes(P.field(k)))
print("TABLES", h.hexdigest())
'''


@pytest.mark.ref
@pytest.mark.parametrize("ice", ["spice_bfr-v2", "spice_ftp-v3m"])
def test_tables_match_reference_real_ice(ice):
    if not U.ref_available() or not os.path.isdir(os.path.join(REFICE, ice)):
        pytest.skip("reference ice directories not present")
    out = _child(_CMP, os.path.join(REFICE, ice), HERE)
    assert "DIFF" not in out, out


def _ppc(args, env, stdin=None):
    from ppc_amd import _native

    e = dict(os.environ)
    for k in ("WFLA", "FWID", "VTHK"):
        e.pop(k, None)
    e.update(env)
    e["DTMN"] = "1"
    r = subprocess.run([_native.ppc_exe()] + args, env=e, input=stdin, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def _rows(text):
    return np.array([[float(v) for v in l.split()] for l in text.strip().splitlines()])


@pytest.mark.parametrize("cfg", ["c1", "c2", "c2x"])
def test_cli_table_dump(cfg):
    """`ppc -`: tilt-corrected layer depth, abs, sca*(1-g), bfr scattering at the middle wavelength bin"""
    from ppc_amd import configs

    env = dict(ENVS[cfg], PPCTABLESDIR=configs.ensure(cfg))
    env.pop("FWID", None)
    got = _rows(_ppc(["-"], env))
    ref = _rows(open(os.path.join(GOLD, "table_%s.txt" % cfg)).read())
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("cfg", ["c2", "c2x"])
def test_cli_zshift(cfg):
    """`ppc =`: depth shift of the tilted layers and the tilt-corrected layer thickness (VTHK=1)"""
    from ppc_amd import configs

    ref_text = open(os.path.join(GOLD, "zshift_%s.txt" % cfg)).read()
    inp = "\n".join(" ".join(l.split()[:3]) for l in ref_text.strip().splitlines()) + "\n"
    got = _rows(_ppc(["="], {"PPCTABLESDIR": configs.ensure(cfg), "VTHK": "1"}, stdin=inp))
    ref = _rows(ref_text)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=2e-4)


@pytest.mark.ref
@pytest.mark.parametrize("ice", ["spice_bfr-v2", "spice_ftp-v3m"])
def test_cli_table_dump_real_ice(ice):
    """`ppc -`, `ppc - x y` and `ppc =` on the reference's own ice directories (1-d and 2-d
    tilt) against the reference program's output (oracle/_ref, run here)"""
    path = os.path.join(REFICE, ice)
    if not U.ref_available() or not os.path.isdir(path):
        pytest.skip("reference ice directories not present")
    for args in (["-"], ["-", "310.5", "-205.25"]):
        e = dict(os.environ)
        for k in ("WFLA", "FWID", "VTHK"):
            e.pop(k, None)
        e.update({"PPCTABLESDIR": path, "DTMN": "1"})
        ref = subprocess.run([U.REF_EXE] + args, env=e, capture_output=True, text=True, timeout=120)
        assert ref.returncode == 0, ref.stderr[-2000:]
        got = _rows(_ppc(args, {"PPCTABLESDIR": path}))
        want = _rows(ref.stdout)
        assert got.shape == want.shape
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-6)
    # `ppc =`: the depth shift of the tilted layers at points across the detector (VTHK=1)
    pts = np.random.default_rng(9).uniform([-700, -700, -600], [700, 700, 600], (300, 3))
    inp = "\n".join("%.3f %.3f %.3f" % tuple(p) for p in pts) + "\n"
    e = dict(os.environ)
    for k in ("WFLA", "FWID"):
        e.pop(k, None)
    e.update({"PPCTABLESDIR": path, "DTMN": "1", "VTHK": "1"})
    ref = subprocess.run([U.REF_EXE, "="], env=e, input=inp, capture_output=True, text=True, timeout=120)
    assert ref.returncode == 0, ref.stderr[-2000:]
    got = _rows(_ppc(["="], {"PPCTABLESDIR": path, "VTHK": "1"}, stdin=inp))
    np.testing.assert_allclose(got, _rows(ref.stdout), rtol=1e-5, atol=2e-4)


@pytest.mark.parametrize("cfg", ["c2", "c2x"])
def test_cli_table_dump_at_position(cfg):
    """`ppc - x y`: the same table with the depth shift taken at a horizontal position"""
    from ppc_amd import configs

    got = _rows(_ppc(["-", "310.5", "-205.25"], {"PPCTABLESDIR": configs.ensure(cfg)}))
    ref = _rows(open(os.path.join(GOLD, "table_%s_xy.txt" % cfg)).read())
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("cfg", ["c2", "c2x"])
def test_cli_zshift_grid(cfg):
    """`ppc _`: the first 25100 lines of the depth-shift grid dump (z, x, y, shift)"""
    from ppc_amd import _native, configs

    ref = _rows(open(os.path.join(GOLD, "grid_%s.txt" % cfg)).read())
    e = dict(os.environ)
    for k in ("WFLA", "FWID", "VTHK"):
        e.pop(k, None)
    e.update({"PPCTABLESDIR": configs.ensure(cfg), "DTMN": "1"})
    p = subprocess.Popen([_native.ppc_exe(), "_"], env=e, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
    try:
        lines = [p.stdout.readline() for _ in range(len(ref))]
    finally:
        p.kill()
        p.wait()
    got = _rows("".join(lines))
    assert got.shape == ref.shape
    np.testing.assert_allclose(got[:, :3], ref[:, :3], rtol=0, atol=0)
    np.testing.assert_allclose(got[:, 3], ref[:, 3], rtol=1e-5, atol=2e-4)


def test_zshift_1d_by_hand():
    """1-d tilt: linear interpolation in distance along 225 deg between the two
    bracketing reference strings and linear in depth between tabulated rows
    (pro.cu:169-191), computed here from the tilt files directly."""
    from ppc_amd import api, configs

    d = configs.ensure("c2")
    par = np.loadtxt(os.path.join(d, "tilt.par"))
    dat = np.loadtxt(os.path.join(d, "tilt.dat"))
    out = _child(r'''
import sys
from ppc_amd import api
D = api.Detector(sys.argv[1])
import numpy as np
pts = np.random.default_rng(3).uniform([-500, -500, -500], [500, 500, 500], (50, 3))
r0 = D.scalar("r0")
for p in pts:
    s, dz = D.zshift(*p)
    print(p[0], p[1], p[2], s, r0)
''', d, env={})
    ang = np.radians(225.0)
    lnx, lny = np.cos(ang), np.sin(ang)
    dist = par[:, 1]
    depth = dat[:, 0]
    zoff = 1948.07
    for line in out.strip().splitlines():
        x, y, z, s, r0 = map(float, line.split())
        nr = lnx * x + lny * y - r0
        j = int(np.searchsorted(dist, nr, side="right"))
        j = min(max(j, 1), len(dist) - 1)
        i = j - 1
        w = (nr - dist[i]) / (dist[j] - dist[i])
        dep = zoff - z
        col = (1 - w) * dat[:, 1 + i] + w * dat[:, 1 + j]
        # rows run from the deepest up; clamp at the ends, extrapolate linearly inside the last cell
        exp = np.interp(dep, depth, col)
        if depth[0] <= dep <= depth[-1]:
            assert abs(s - exp) < 2e-3 * (1 + abs(exp)), (x, y, z, s, exp)


def test_mwc_multipliers_reproduce_reference_table():
    """rnd.txt of the ice directories is the descending sequence of safe-prime
    MWC multipliers below 2^32; the search reproduces all 131072 entries."""
    from ppc_amd import configs

    g = json.load(open(os.path.join(GOLD, "mwc_rnd.json")))
    vals = configs.mwc_multipliers(g["count"])
    assert list(vals[:16]) == g["first"]
    assert list(vals[-4:]) == g["last"]
    text = "\n".join(str(int(v)) for v in vals) + "\n"
    assert hashlib.sha256(text.encode()).hexdigest() == g["sha256"]
    # definition: a*2^32-1 and (a*2^32-2)/2 both prime (check a few with sympy-free Fermat tests)
    for a in vals[:5].tolist() + vals[-3:].tolist():
        p = (a << 32) - 1
        for base in (2, 3, 5, 7, 11):
            assert pow(base, p - 1, p) == 1
            q = (p - 1) // 2
            assert pow(base, q - 1, q) == 1
