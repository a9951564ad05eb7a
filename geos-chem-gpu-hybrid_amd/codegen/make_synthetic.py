"""Write the synthetic mechanism IRs under mechanisms/.

* stress500.json  the stress mechanism (500 species, 1500 reactions,
                  ~3% LU fill), synthetic.stress_mechanism(seed=4)
* auto1.json      one species, X -> 2X: with rate 1/8 and H = 16 the
                  prepared diagonal -J + 1/(0.5 H) is exactly zero, which
                  drives KppDecomp's singular-matrix retry on purpose
* quad1.json      one species, 2X -> (nothing) and F -> X: a scalar
                  nonlinear problem whose step control rejects attempts
                  before and after the first accepted step
"""
import importlib.util
import json
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parents[1]


def _synthetic():
    spec = importlib.util.spec_from_file_location("gc_synthetic", PKG / "synthetic.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gc_synthetic"] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    s = _synthetic()
    out = PKG / "mechanisms"
    irs = {
        "stress500": s.stress_mechanism(),
        "auto1": s.build_ir("auto1", 1, 1, [s.Reaction({0: 1}, {0: 2.0})], reorder=False),
        "quad1": s.build_ir("quad1", 1, 1, [s.Reaction({0: 2}, {}),
                                            s.Reaction({}, {0: 1.0}, {0: 1})], reorder=False),
    }
    for name, ir in irs.items():
        (out / f"{name}.json").write_text(json.dumps(ir, separators=(",", ":")))
        print(name, ir["nvar"], ir["nreact"], ir["nnz"])


if __name__ == "__main__":
    sys.exit(main())
