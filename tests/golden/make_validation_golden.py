"""Extract the FP64 validation numbers the reference committed (results/validation.json) into
tests/golden/validation_golden.json, a data fixture for the oracle tests.

Run where the reference tree is checked out (the GPU tests never read it):
    python tests/golden/make_validation_golden.py /root/reference/results/validation.json
"""

import json
import sys
from pathlib import Path


def main(src: str) -> None:
    d = json.loads(Path(src).read_text())
    keep_run = ("n", "h", "dt", "r", "steps", "iterations", "err", "err_exact",
                "err_semidiscrete", "err_discrete")
    out = {"source": "reference results/validation.json (fp64, torch-eager engine)"}
    for k in ("spatial", "temporal", "laplace", "poisson"):
        out[k] = [{"dims": x["dims"], "bc": x.get("bc", "dirichlet"), "order": x["order"],
                   **{kk: x[kk] for kk in ("r", "t_final", "tol", "n") if kk in x},
                   "runs": [{kk: r[kk] for kk in keep_run if kk in r} for r in x["runs"]]}
                  for x in d[k]]
    out["laplace_quadratic"] = [{kk: x[kk] for kk in ("dims", "n", "iterations", "err")}
                                for x in d["laplace_quadratic"]]
    Path(__file__).with_name("validation_golden.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/results/validation.json")
