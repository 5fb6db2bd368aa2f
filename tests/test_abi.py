"""The C ABI: the library loads, exports every entry point of include/cfx.h,
and the ctypes mirrors of the descriptor structs have the C layout.  No
compute call is made, so this runs without a GPU."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from combustioncodegpu_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cfx.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cfx_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_step_api():
    names = declared_functions()
    for required in ("cfx_create", "cfx_destroy", "cfx_rhs", "cfx_time_step", "cfx_update", "cfx_apply_bc",
                     "cfx_exchange", "cfx_update_properties", "cfx_residual", "cfx_step", "cfx_download",
                     "cfx_upload_primitive", "cfx_init_conservative"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from combustioncodegpu_amd import lib
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libcfx.so not built")
    so = C.CDLL(lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(so, n)]
    assert not missing, missing
    so.cfx_version.restype = C.c_int
    assert so.cfx_version() == 1


def test_loader_fails_loudly_without_library(monkeypatch, tmp_path):
    from combustioncodegpu_amd import lib
    monkeypatch.setattr(lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(lib, "_lib", None)
    with pytest.raises(RuntimeError):
        lib.load()


def test_struct_layouts_match_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "cfx.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(cfx_bc), sizeof(cfx_inner), sizeof(cfx_block),
         sizeof(cfx_inflow), sizeof(cfx_wall), sizeof(cfx_params), sizeof(cfx_species));
  printf("%zu %zu %zu\n", offsetof(cfx_params, cfl), offsetof(cfx_params, inflow), offsetof(cfx_block, inner));
  printf("%zu %zu %zu %zu\n", sizeof(cfx_mlp_config), sizeof(cfx_mlp_report), offsetof(cfx_mlp_config, seed),
         offsetof(cfx_mlp_report, n_params));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(v) for v in out]
    mine = [C.sizeof(abi.CfxBC), C.sizeof(abi.CfxInner), C.sizeof(abi.CfxBlock), C.sizeof(abi.CfxInflow),
            C.sizeof(abi.CfxWall), C.sizeof(abi.CfxParams), C.sizeof(abi.CfxSpecies),
            abi.CfxParams.cfl.offset, abi.CfxParams.inflow.offset, abi.CfxBlock.inner.offset,
            C.sizeof(abi.CfxMlpConfig), C.sizeof(abi.CfxMlpReport), abi.CfxMlpConfig.seed.offset,
            abi.CfxMlpReport.n_params.offset]
    assert sizes == mine
