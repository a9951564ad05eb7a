"""Write tests/golden/fullchem_checkforms.json.gz from the reference checkout.

The fixture holds the reference's *independent* forms of the fullchem
mechanism, parsed as text by codegen/kpp_ingest.py, so the tests can pin the
mechanism IR (and with it the oracle and the generated kernels) against them:

* Fun (aggregate Vdot sums)       gckpp_Function.F90:47-2150
* CalcStoichNum                   gckpp_Function.F90:7271-12790
* Jac_SP_Vec, JacTR_SP_Vec        gckpp_Jacobian.F90:20903-22738
* KppSolveTR statement order      gckpp_LinearAlgebra.F90:2327-3701

Run here (where the reference is mounted):
    python tests/golden/make_checkforms.py /root/reference
"""
import gzip
import importlib.util
import json
import sys
from pathlib import Path

REPO = Path(__file__).resolve().parents[2]
spec = importlib.util.spec_from_file_location(
    "kpp_ingest", REPO / "geos-chem-gpu-hybrid_amd" / "codegen" / "kpp_ingest.py")
kpp_ingest = importlib.util.module_from_spec(spec)
spec.loader.exec_module(kpp_ingest)

if __name__ == "__main__":
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    forms = kpp_ingest.ingest_checkforms(ref / "src/geos-chem/KPP/fullchem")
    out = Path(__file__).resolve().parent / "fullchem_checkforms.json.gz"
    with gzip.open(out, "wt") as f:
        json.dump(forms, f, separators=(",", ":"))
    print(out, out.stat().st_size, "bytes")
