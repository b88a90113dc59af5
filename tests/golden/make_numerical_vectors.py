"""Extract the complex-step derivative vectors of the reference's numerical
MCSE tests (rust/crates/vckss-core/tests/numerical_mc.rs:118-777) into
tests/golden/numerical_mc_vectors.json: for each case the moment means u,
the probe count, the control, and the expected observation gradient, db and
dv.  Reads the reference test file as text; run in the build container:

    python tests/golden/make_numerical_vectors.py /root/reference
"""
import json
import os
import re
import sys


def main(ref):
    path = os.path.join(ref, "rust", "crates", "vckss-core", "tests", "numerical_mc.rs")
    text = open(path).read()
    start = text.index("fn independent_complex_step_derivatives")
    body = text[start:]
    num = r"(-?[0-9.]+(?:e-?[0-9]+)?)_f64"
    cases = []
    for block in body.split("let u = [")[1:]:
        def arr(name, src):
            m = re.search(name + r" = \[(.*?)\];", src, re.S)
            return [float(x) for x in re.findall(num, m.group(1))]
        u = [float(x) for x in re.findall(num, block.split("];")[0])]
        r = int(re.search(r"FiniteDerivative::new\(u, ([0-9]+)\)", block).group(1))
        control = float(re.search(r"observation\(([0-9.e-]+)\)", block).group(1))
        cases.append(dict(u=u, r=r, control=control, observation=arr("let expected", block),
                          db=arr("let expected_db", block), dv=arr("let expected_dv", block)))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "numerical_mc_vectors.json")
    with open(out, "w") as fh:
        json.dump(dict(source="rust/crates/vckss-core/tests/numerical_mc.rs:118-777", cases=cases), fh, indent=1)
    print(f"{len(cases)} cases -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
