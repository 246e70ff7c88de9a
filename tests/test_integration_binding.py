"""INTEGRATION.md's reference-side binding compiles against the reference's own
headers: the C++ block's three parts (the include, the SolverClass members and
the member definitions) are placed where their markers say, in a scratch copy of
VCP3D_CurviLinear/include, and the definitions are compiled with g++
(-fsyntax-only) together with include/vcp3d_amd.h. Needs the reference sources
(present where the reference is checked out) and g++; skipped otherwise."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = "/root/reference/VCP3D_CurviLinear/include"
MARK_TOP = "// SolverClass.hpp, with its includes:"
MARK_MEMBERS = "// SolverClass.hpp, private members of SolverClass:"
MARK_CPP = "// SolverClass.cpp:"


def binding_block() -> str:
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    found = [b for b in blocks if MARK_TOP in b]
    assert len(found) == 1, "one C++ binding block with the include marker"
    return found[0]


def split_block(block: str):
    top = block.index(MARK_TOP)
    mem = block.index(MARK_MEMBERS)
    cpp = block.index(MARK_CPP)
    assert top < mem < cpp
    return block[top:mem], block[mem:cpp], block[cpp:]


def test_binding_block_is_well_formed():
    head, members, cpp = split_block(binding_block())
    assert '#include "vcp3d_amd.h"' in head
    assert "vcp_ctx *mGpu" in members
    for name in ("SolverClass::CreateGpuContext", "SolverClass::Residual",
                 "SolverClass::ExecuteTimeSynchronizationStep"):
        assert name in cpp


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE) or shutil.which("g++") is None,
                    reason="needs the reference headers and g++")
def test_binding_compiles_against_reference_headers(tmp_path):
    head, members, cpp = split_block(binding_block())
    inc = tmp_path / "include"
    shutil.copytree(REF_INCLUDE, inc)
    hdr = inc / "SolverClass.hpp"
    text = hdr.read_text()
    # the include goes beside the header's own includes, the members into the
    # class's private section (its first "private:")
    anchor = '#include "ElementClass.hpp"\n'
    assert text.count(anchor) == 1 and text.count("private:") >= 1
    text = text.replace(anchor, anchor + head, 1)
    text = text.replace("private:\n", "private:\n" + members, 1)
    hdr.write_text(text)
    src = tmp_path / "binding.cpp"
    src.write_text('#include "VCP3D_CurviLinear.hpp"\n' + cpp)
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-w", f"-I{inc}", f"-I{os.path.join(ROOT, 'include')}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
