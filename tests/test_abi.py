"""CPU checks of the drop-in boundary: libquac.so builds, loads, exports every
function that include/quac.h, include/quac_hip.h and include/quac_legacy.h
declare (and the legacy API's globals), and the product path fails loudly (no
fallback) when no HIP device is present."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "quac_amd", "libquac.so")


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    txt = "\n".join(l for l in txt.splitlines()
                    if not l.strip().startswith("#") and l.strip() not in ('extern "C" {', "}"))
    # drop struct/enum bodies
    while True:
        t2 = re.sub(r"\{[^{}]*\}", "", txt)
        if t2 == txt:
            break
        txt = t2
    names = []
    for stmt in txt.split(";"):
        s = " ".join(stmt.split())
        if not s or s.startswith("typedef") or "(*" in s.split("(")[0] or s.startswith("extern \"C\""):
            continue
        s = s.replace('extern "C"', "").strip()
        m = re.match(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", s)
        if m:
            names.append(m.group(1))
    return sorted(set(names))


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], stdout=subprocess.PIPE, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


@pytest.mark.parametrize("header", ["quac.h", "quac_hip.h", "quac_legacy.h"])
def test_every_declared_symbol_is_exported(header):
    assert os.path.exists(LIB), "build libquac.so first (python -m quac_amd.build)"
    names = declared(header)
    assert len(names) > 20
    missing = [n for n in names if n not in exported()]
    assert not missing, missing


def test_legacy_globals_are_exported():
    """operators.h:60-66 / quac.c:11-12: nid, np, total_levels, num_subsystems,
    subsystem_list are data symbols of the library"""
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], stdout=subprocess.PIPE, text=True, check=True).stdout
    data = {l.split()[-1] for l in out.splitlines() if l.split()[1] in ("B", "D")}
    for n in ("nid", "np", "total_levels", "num_subsystems", "subsystem_list"):
        assert n in data, n


def test_legacy_programs_compile_against_the_headers(tmp_path):
    """the legacy C test programs build against include/ and link with libquac.so"""
    for prog in ("legacy_jc", "legacy_vec", "legacy_steady", "legacy_circuit", "legacy_dm"):
        subprocess.check_call(["gcc", "-O2", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                               os.path.join(ROOT, "tests", "c", prog + ".c"), "-o", str(tmp_path / prog),
                               "-L", os.path.dirname(LIB), "-lquac", "-Wl,-rpath," + os.path.dirname(LIB), "-lm"])


def test_loads_and_fails_loudly_without_device():
    import quac_amd as qa
    L = qa.lib()
    if qa.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(qa.QuacError):
        qa.System()


def test_td_config_validates_its_switches():
    """qh_td_config needs no device: out-of-range switches are errors, valid
    ones are taken (and the defaults restored)"""
    import quac_amd as qa
    qa.lib()
    for bad in ((-1, 1, 0), (8193, 1, 0), (4096, 2, 0), (4096, 1, 3), (4096, 1, 32)):
        with pytest.raises(qa.QuacError):
            qa.td_config(*bad)
    for ok in ((0, 0, 0), (8192, 1, 1), (4096, 1, 4), (4096, 0, 16)):
        qa.td_config(*ok)
    qa.td_config()


def test_exit_semantics_match_reference_in_a_child():
    """Misuse prints ERROR! and exits (reference: exit(0) in create_op_sys after
    construct_matrix, qsystem.c:158-162) -- default error mode, run in a child."""
    code = r'''
import ctypes, sys
L = ctypes.CDLL(sys.argv[1])
L.quac_set_error_mode(0)
L.QuaC_initialize.argtypes = [ctypes.c_int, ctypes.c_void_p]
L.QuaC_initialize(0, None)
'''
    r = subprocess.run(["python", "-c", code, LIB], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # without a device QuaC_initialize reports the missing device and exits(9)
    import quac_amd as qa
    if qa.device_count() == 0:
        assert r.returncode == 9 and "ERROR!" in r.stdout
