"""Extract the all-electron C2 wavefunction and VMC results of CHAMP's shipped
transcript (examples/c2_ktt_hfj_vmc_emin_j_opt3_lin.{inp,out}; old-format
input, Slater basis, isc = 2 Jastrow) into tests/golden/c2_ktt_transcript.json.

Reads the reference's example files as text (data); nothing is executed.
Each record: the Jastrow parameters of one VMC run (iteration 1: the
input's own; iteration k > 1: the "New wave function" block printed after
iteration k - 1) and that run's "total E" line (energy, error, sigma, sigma
error), the number of passes and the acceptance ratio.
"""
import json
import os
import re

REF = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))
STEM = "c2_ktt_hfj_vmc_emin_j_opt3_lin"


def _nums(line):
    out = []
    for tok in line.split():
        try:
            out.append(float(tok))
        except ValueError:
            break
    return out


def _acceptance(out, j):
    """The run's "acceptance" line (the accelerated Metropolis mover, imetro
    = 6, CHAMP's default) after its "total E" line."""
    while not out[j].startswith("acceptance"):
        j += 1
    return float(out[j].split()[1])


def main():
    inp = open(os.path.join(REF, STEM + ".inp")).read()
    out = open(os.path.join(REF, STEM + ".out")).read().splitlines()
    runs, params = [], None
    for i, ln in enumerate(out):
        m = re.match(r"Beginning optimization iteration #\s*(\d+)", ln)
        if m:
            it = int(m.group(1))
            j = i
            while not out[j].startswith("Final results after"):
                j += 1
            passes = float(out[j].split()[3])
            while not out[j].startswith("total E ="):
                j += 1
            v = out[j].split()
            runs.append(dict(iteration=it, params=params, energy=float(v[3]), error=float(v[5]),
                             sigma=float(v[8]), sigma_error=float(v[10]), passes=passes,
                             acceptance=_acceptance(out, j)))
        if ln.startswith("New wave function:"):
            j = i + 1
            while out[j].strip() != "Jastrow parameters:":
                j += 1
            rows = [_nums(out[j + 1 + k]) for k in range(3)]
            params = dict(a=rows[0], b=rows[1], c=rows[2])
    json.dump(dict(input=inp, runs=runs), open(os.path.join(HERE, "c2_ktt_transcript.json"), "w"),
              indent=1)
    for r in runs:
        print(r["iteration"], r["energy"], r["error"], r["sigma"], r["passes"],
              None if r["params"] is None else [len(r["params"][k]) for k in "abc"])


if __name__ == "__main__":
    main()
