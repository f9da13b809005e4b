"""C-ABI surface: the library loads, exports every symbol include/v2c.h
declares, and its tables agree with the Python mirror (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import pytest

from velo2cycles_amd import abi

HEADER = abi.HEADER_PATH


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(v2c_\w+)\s*\(", text)))


def test_header_enum_matches_python():
    assert abi.header_enum_names() == list(abi.NODAL_FIELDS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(abi.LIB_PATH), "libv2c.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (v2c_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_library_loads_and_tables_agree():
    lib = abi.load_lib()
    for i, name in enumerate(abi.NODAL_FIELDS):
        assert lib.v2c_nodal_name(i).decode() == name
    assert lib.v2c_nodal_name(abi.NF) is None
    p = abi.Params()
    lib.v2c_params_default(ctypes.byref(p))
    q = abi.default_params()
    for k, _ in abi.Params._fields_:
        if k == "xstpx":
            continue
        assert getattr(p, k) == getattr(q, k), k


def test_struct_layout_matches_c():
    """sizeof of the ABI structs seen by a C compiler equals the ctypes mirror."""
    src = ('#include "v2c.h"\n#include <stdio.h>\nint main(void){printf("%zu %zu %zu %zu\\n",'
           'sizeof(v2c_grid),sizeof(v2c_params),sizeof(v2c_materials),sizeof(v2c_markers));return 0;}')
    exe = "/tmp/v2c_sizes"
    subprocess.run(["gcc", "-x", "c", "-I", os.path.dirname(HEADER), "-o", exe, "-"], input=src,
                   text=True, check=True)
    sizes = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    assert sizes == [ctypes.sizeof(abi.Grid), ctypes.sizeof(abi.Params),
                     ctypes.sizeof(abi.Materials), ctypes.sizeof(abi.Markers)]


def test_result_structs_match_c():
    """sizeof and field offsets of the result structs (v2c_sticky_stats,
    v2c_timestep_params / _result) equal the ctypes mirrors."""
    names = [n for n, _ in abi.StickyStats._fields_]
    offs = ",".join(f"offsetof(v2c_sticky_stats,{n})" for n in names)
    src = ('#include "v2c.h"\n#include <stdio.h>\n#include <stddef.h>\n'
           'int main(void){size_t o[]={' + offs + '};'
           'printf("%zu %zu %zu",sizeof(v2c_sticky_stats),sizeof(v2c_timestep_params),sizeof(v2c_timestep_result));'
           'for(unsigned i=0;i<sizeof o/sizeof o[0];i++)printf(" %zu",o[i]);printf("\\n");return 0;}')
    exe = "/tmp/v2c_result_sizes"
    subprocess.run(["gcc", "-x", "c", "-I", os.path.dirname(HEADER), "-o", exe, "-"], input=src,
                   text=True, check=True)
    got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    assert got[:3] == [ctypes.sizeof(abi.StickyStats), ctypes.sizeof(abi.TimestepParams),
                       ctypes.sizeof(abi.TimestepResult)]
    assert got[3:] == [getattr(abi.StickyStats, n).offset for n in names]


def test_prn_struct_matches_c():
    """The ctypes mirror of v2c_prn has the C struct's size and field offsets."""
    from velo2cycles_amd.prn import Prn
    fields = ["marknum", "pkf", "rocknum", "bondnum", "cycle", "markim", "rock", "node", "bondm", "gx",
              "bondn", "timesave", "mark", "t"]
    body = ",".join(f"(size_t)offsetof(v2c_prn, {f})" for f in fields)
    fmt = " ".join(["%zu"] * (len(fields) + 1))
    src = ('#include "v2c.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(void){printf("' + fmt +
           '\\n", sizeof(v2c_prn),' + body + ');return 0;}')
    exe = "/tmp/v2c_prn_layout"
    subprocess.run(["gcc", "-x", "c", "-I", os.path.dirname(HEADER), "-o", exe, "-"], input=src,
                   text=True, check=True)
    got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    assert got == [ctypes.sizeof(Prn)] + [getattr(Prn, f).offset for f in fields]


def test_dropin_binding_compiles_cleanly():
    """The reference-side binding is plain C99 against the header: no warnings."""
    src = os.path.join(os.path.dirname(abi.LIB_PATH), "csrc", "dropin", "v2c_dropin.c")
    subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter", "-fsyntax-only",
                    src], check=True)


def test_create_without_gpu_fails_cleanly():
    """With no visible device v2c_create returns ENOGPU, never crashes."""
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is visible")
    from velo2cycles_amd import synth
    cfg = synth.config_grid(11, 9, 10e3, 8e3, 2)
    with pytest.raises(abi.V2CError, match="ENOGPU|EHIP"):
        abi.Context(cfg.grid, cfg.params, cfg.mats, capacity=cfg.n)
