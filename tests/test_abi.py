"""The C ABI library loads (no GPU needed), exports every function include/lg.h declares, and
the ctypes mirror agrees with the C layout of every struct."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from light_garden_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lg.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lg_[a-z0-9_]+)\s*\(", src)))


def test_header_functions_exported(lglib):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lglib, n)]
    assert not missing, missing


def test_ctypes_mirror_covers_header():
    assert set(declared_functions()) == set(abi.FUNCTIONS)


def test_version_and_error_without_device(lglib):
    assert lglib.lg_version() == 1
    # calls that need no device still validate their arguments
    assert lglib.lg_scene_create(None, None) == abi.LG_E_INVALID
    assert b"NULL" in lglib.lg_last_error()


def test_struct_layout_matches_c():
    """Compile a probe against lg.h with gcc and compare sizeof/offsetof with ctypes."""
    fields = {
        "lg_shape": abi.LgShape, "lg_object": abi.LgObject, "lg_light": abi.LgLight,
        "lg_params": abi.LgParams, "lg_view": abi.LgView, "lg_stage_time": abi.LgStageTime,
        "lg_tuning": abi.LgTuning, "lg_server_config": abi.LgServerConfig, "lg_request": abi.LgRequest,
        "lg_result": abi.LgResult, "lg_server_stats": abi.LgServerStats,
        "lg_mod_rem_color": abi.LgModRemColor, "lg_string_mod": abi.LgStringMod,
        "lg_blend_component": abi.LgBlendComponent, "lg_blend": abi.LgBlend,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){"]
    for cname, cls in fields.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(l.rsplit(" ", 1) for l in out.strip().splitlines())
    for cname, cls in fields.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f in cls._fields_:
            assert int(got["%s.%s" % (cname, f[0])]) == getattr(cls, f[0]).offset, (cname, f[0])
