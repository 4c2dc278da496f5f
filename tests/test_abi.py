"""The C ABI without a GPU: the library loads, exports every function the
headers declare, the ctypes mirror matches the C struct layouts, the
drop-in's SAMPLE_MOLECULE restatement matches the reference's compiled
layout, and GPU entry points fail loudly (no CPU fallback)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from mspms2_amd import abi
from mspms2_amd.engine import EXPORTS, EngineError, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def _declared(header):
    with open(os.path.join(INC, header)) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mspms_\w+|frclong|frcshort)\s*\(",
                          txt, flags=re.M))


def test_every_declared_function_is_exported():
    names = _declared("mspms_gpu.h")
    assert len(names) >= 20, names
    L = lib()
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(EXPORTS) <= names


def test_dropin_archive_exports_frclong_frcshort():
    a = os.path.join(ROOT, "mspms2_amd", "libmspms_dropin.a")
    assert os.path.exists(a), "build first (python -m mspms2_amd.build)"
    out = subprocess.run(["nm", "-g", "--defined-only", a], capture_output=True,
                         text=True).stdout
    for n in _declared("mspms_dropin.h"):
        assert re.search(rf"\bT {n}\b", out), n


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "mspms_gpu.h"
#define S(t) printf(#t " %zu\n", sizeof(t));
#define O(t, f) printf(#t "." #f " %zu\n", offsetof(t, f));
int main(void) {
  S(mspms_species) S(mspms_params) S(mspms_long_terms) S(mspms_short_terms)
  S(mspms_sf_grid) S(mspms_param_grad) S(mspms_tune_result) S(mspms_nvt_state)
  O(mspms_tune_result, cand_ms) O(mspms_tune_result, cand_wpc)
  O(mspms_nvt_state, pss) O(mspms_nvt_state, tinst) O(mspms_nvt_state, nfree)
  S(mspms_npt_state) O(mspms_npt_state, pinst) O(mspms_npt_state, boxl)
  O(mspms_npt_state, boxv) O(mspms_npt_state, lj_tail)
  O(mspms_species, nnbp) O(mspms_species, imp_para) O(mspms_species, dih_unique)
  O(mspms_params, kappa) O(mspms_params, coulomb_prefactor) O(mspms_params, wolffcon2)
  O(mspms_params, ntube) O(mspms_params, solid_natom) O(mspms_params, grid)
  O(mspms_long_terms, virial_inter) O(mspms_long_terms, elec_pair_hash)
  O(mspms_sf_grid, nodes) O(mspms_sf_grid, zperiod)
  S(mspms_sfgridgen_spec) O(mspms_sfgridgen_spec, isLJswitch) O(mspms_sfgridgen_spec, uclx)
  O(mspms_sfgridgen_spec, s_epsilon) O(mspms_sfgridgen_spec, sz)
  return 0;
}
"""


def test_ctypes_layout_matches_c():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "p.c")
        with open(src, "w") as f:
            f.write(PROBE)
        exe = os.path.join(tmp, "p")
        subprocess.run(["gcc", "-I", INC, src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = dict(ln.rsplit(" ", 1) for ln in out.strip().splitlines())
    py = {"mspms_species": abi.Species, "mspms_params": abi.Params,
          "mspms_long_terms": abi.LongTerms, "mspms_short_terms": abi.ShortTerms,
          "mspms_sf_grid": abi.SFGrid, "mspms_param_grad": abi.ParamGrad,
          "mspms_tune_result": abi.TuneResult, "mspms_nvt_state": abi.NvtState,
          "mspms_npt_state": abi.NptState, "mspms_sfgridgen_spec": abi.SfGridGenSpec}
    for k, v in got.items():
        if "." in k:
            t, f = k.split(".")
            assert getattr(py[t], f).offset == int(v), k
        else:
            assert C.sizeof(py[k]) == int(v), k


def test_dropin_sample_molecule_layout_matches_reference():
    """mspms2_amd/csrc/ref_layout.h restates SAMPLE_MOLECULE
    (src/molevars.h:40-73); oracle/build_ref.sh records the reference's
    compiled offsets in oracle/_ref/sample_molecule_layout.txt."""
    ref = os.path.join(ROOT, "oracle", "_ref", "sample_molecule_layout.txt")
    if not os.path.exists(ref):
        pytest.skip("reference layout not recorded (built only where the reference is)")
    want = dict(ln.split() for ln in open(ref).read().strip().splitlines())
    fields = [k for k in want if k != "sizeof"]
    body = "\n".join(f'  printf("{k} %zu\\n", offsetof(ref_sample_molecule, {k}));'
                     for k in fields)
    src = ('#include <stdio.h>\n#include <stddef.h>\n#include "ref_layout.h"\n'
           'int main(void){ printf("sizeof %zu\\n", sizeof(ref_sample_molecule));\n'
           + body + "\n return 0; }\n")
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "l.c")
        with open(p, "w") as f:
            f.write(src)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "mspms2_amd", "csrc"), p,
                        "-o", os.path.join(tmp, "l")], check=True)
        out = subprocess.run([os.path.join(tmp, "l")], capture_output=True, text=True,
                             check=True).stdout
    got = dict(ln.split() for ln in out.strip().splitlines())
    assert got == want


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU failure path")
def test_no_gpu_fails_loudly():
    from mspms2_amd import configs
    from mspms2_amd.engine import Engine
    from tests.util import system_of
    s = system_of(configs.lj_fcc(ncell=2, rc=1.2))
    with pytest.raises(EngineError, match="EIO"):
        Engine(s)


def test_derive_params_rejects_large_cutoff():
    p = abi.Params(sigma_base=1.0, epsilon_base=1.0, boxlx=10.0, boxly=10.0, boxlz=10.0,
                   rcuton=6.0, rcutoff=6.0, rcutoffelec=6.0)
    assert lib().mspms_derive_params(C.byref(p)) == abi.MSPMS_ERANGE
    assert b"half the box" in lib().mspms_last_error()
