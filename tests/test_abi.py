"""The C-ABI library loads on a machine without a GPU and exports every
entry point include/ampe_hip.h declares (no compute calls)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ampe_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ampe_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "ampe_phase_rhs" in syms and "ampe_sim_advance" in syms
    assert len(syms) >= 25


def test_library_exports_every_declared_symbol():
    from ampe_amd import hip as H
    lib = C.CDLL(H.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    from ampe_amd import hip as H
    assert set(declared_symbols()) == set(H.EXPORTED)


def test_struct_sizes_match_header():
    """ctypes mirrors have the C layout (compiled probe of sizeof/offsetof)."""
    import subprocess
    import tempfile
    from ampe_amd import params as P
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "ampe_hip.h"
int main(void){
 printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(ampe_model_params),
        sizeof(ampe_solver_params), sizeof(ampe_stats), sizeof(ampe_box),
        offsetof(ampe_model_params, calphad), offsetof(ampe_model_params, newton_tol),
        sizeof(ampe_elliptic_op), sizeof(ampe_calphad_gfun), sizeof(ampe_calphad_tdep),
        offsetof(ampe_calphad_tdep, L), sizeof(ampe_train_params), sizeof(ampe_kks_batch_stats));
 return 0;}
'''
    with tempfile.TemporaryDirectory() as d:
        cf = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(cf, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), cf, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    from ampe_amd.hip import EllipticOp
    got = [C.sizeof(P.ModelParams), C.sizeof(P.SolverParams), C.sizeof(P.Stats),
           C.sizeof(P.Box), P.ModelParams.calphad.offset, P.ModelParams.newton_tol.offset,
           C.sizeof(EllipticOp), C.sizeof(P.CalphadGfun), C.sizeof(P.CalphadTdep),
           P.CalphadTdep.L.offset, C.sizeof(P.TrainParams), C.sizeof(P.KksBatchStats)]
    assert [int(x) for x in out] == got
