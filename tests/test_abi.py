"""The C ABI: the library loads, exports every symbol the header declares,
and the ctypes mirrors have the C layouts (no compute calls here)."""

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from jaxrens_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "jaxrens_amd.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(jr_\w+)\(",
                                 text, flags=re.M)))


def test_header_and_binding_agree():
    names = [n for n, _, _ in _lib.SIGNATURES]
    assert sorted(names) == declared_symbols()


def test_library_exports_every_symbol():
    lib = _lib.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(jr_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "jaxrens_amd.h"\n'
        "int main(void){\n"
        ' printf("%zu %zu %zu %zu %zu\\n", sizeof(jr_pot), sizeof(jr_move),'
        " sizeof(jr_config), sizeof(jr_tape), sizeof(jr_iter_out));\n"
        ' printf("%zu %zu %zu\\n", offsetof(jr_config, seed), offsetof(jr_config, pot),'
        " offsetof(jr_config, moves));\n"
        ' printf("%zu %zu %zu\\n", sizeof(jr_mpnn), offsetof(jr_mpnn, rc),'
        " offsetof(jr_mpnn, emb));\n"
        ' printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(jr_learned),'
        " offsetof(jr_learned, lr), offsetof(jr_learned, sur_params),"
        " sizeof(jr_learned_stats), offsetof(jr_learned_stats, heldout_rmse),"
        " offsetof(jr_config, mpnn), offsetof(jr_config, learned));\n"
        " return 0; }\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True,
                           check=True).stdout.split("\n")
    sizes = [int(x) for x in lines[0].split()]
    offs = [int(x) for x in lines[1].split()]
    assert sizes == [C.sizeof(_lib.JrPot), C.sizeof(_lib.JrMove),
                     C.sizeof(_lib.JrConfig), C.sizeof(_lib.JrTape),
                     C.sizeof(_lib.JrIterOut)]
    assert offs == [_lib.JrConfig.seed.offset, _lib.JrConfig.pot.offset,
                    _lib.JrConfig.moves.offset]
    mp = [int(x) for x in lines[2].split()]
    assert mp == [C.sizeof(_lib.JrMPNN), _lib.JrMPNN.rc.offset, _lib.JrMPNN.emb.offset]
    lw = [int(x) for x in lines[3].split()]
    assert lw == [C.sizeof(_lib.JrLearned), _lib.JrLearned.lr.offset,
                  _lib.JrLearned.sur_params.offset, C.sizeof(_lib.JrLearnedStats),
                  _lib.JrLearnedStats.heldout_rmse.offset, _lib.JrConfig.mpnn.offset,
                  _lib.JrConfig.learned.offset]


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        _lib.load_library(tmp_path / "nope.so")
