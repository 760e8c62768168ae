"""The C ABI: include/gacelle.h compiled by gcc agrees with the ctypes
mirror field by field, and the built library exports every declared
entry point (no compute calls: these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from gacelle_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gacelle.h")
STRUCTS = ["gc_model", "gc_askadam_cfg", "gc_volume", "gc_mcmc_cfg", "gc_mcmc_io"]


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|char\s*\*|const char\s*\*)\s*\*?\s*(gc_\w+)\s*\(", text,
                                 re.M)))


def test_header_declares_exactly_the_mirrored_functions():
    assert _declared_functions() == sorted(_abi.SIGNATURES)


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("abi")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "gacelle.h"', "int main(void){"]
    for s in STRUCTS:
        lines.append(f'printf("{s} sizeof %zu\\n", sizeof({s}));')
        for name, _ in getattr(_abi, s)._fields_:
            cname = "lambda" if name == "lambda_" else name
            lines.append(f'printf("{s} {name} %zu\\n", offsetof({s}, {cname}));')
    for k, v in vars(_abi).items():
        if k.startswith("GC_") and isinstance(v, int) and k not in ("GC_MAX_VOL",):
            lines.append(f'printf("const {k} %lld\\n", (long long)({k}));')
    lines.append("return 0;}")
    src = d / "layout.c"
    src.write_text("\n".join(lines))
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}


@pytest.mark.parametrize("struct", STRUCTS)
def test_struct_layout_matches_header(c_layout, struct):
    cls = getattr(_abi, struct)
    assert c_layout[(struct, "sizeof")] == ctypes.sizeof(cls)
    for name, _ in cls._fields_:
        assert c_layout[(struct, name)] == getattr(cls, name).offset, name


def test_constants_match_header(c_layout):
    consts = {k[1]: v for k, v in c_layout.items() if k[0] == "const"}
    assert len(consts) > 40
    for k, v in consts.items():
        assert getattr(_abi, k) == v, k


def test_header_compiles_as_cplusplus(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "gacelle.h"\nint main(){ return gc_version == nullptr; }\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.dirname(HEADER), str(src)], check=True)


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], check=True, capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\sT\s(gc_\w+)$", out, re.M))
    assert set(_declared_functions()) <= exported
    # nothing beyond the declared C ABI leaks with the gc_ prefix
    assert exported == set(_declared_functions())


@pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="HIP library not built")
def test_library_loads_and_reports_abi_version():
    """dlopen resolves every symbol with its ctypes signature; gc_version is
    the only call that touches no device."""
    lib = _abi.load_library()
    assert lib.gc_version() == _abi.GC_ABI_VERSION
    for name in _abi.SIGNATURES:
        assert getattr(lib, name).argtypes == _abi.SIGNATURES[name][0]


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_abi.GacelleError, match="not built"):
        _abi.load_library(str(tmp_path / "libgacelle_hip.so"))
