"""Extract the Be wavefunction parameters and VMC results of CHAMP's shipped
transcript (examples/be_vb1_jcas24_vmc_j_opt_lin.{inp,out}) into
tests/golden/be_transcript.json.

Reads the reference's example files as text (data); nothing is executed.
The first "sites:" line (the starting configuration from the seed) is kept.
Each record: the Jastrow parameters CHAMP used for one VMC run ("Wave
function for next iteration #k") and that run's "total E" line (energy,
error, sigma, sigma error), the number of passes and the acceptance ratio.
"""
import json
import os
import re

REF = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))


def _acceptance(out, j):
    """The run's "acceptance" line (the accelerated Metropolis mover, imetro
    = 6, CHAMP's default) after its "total E" line."""
    while not out[j].startswith("acceptance"):
        j += 1
    return float(out[j].split()[1])


def _sites(out):
    """The first configuration CHAMP generated with its sites routine (the
    "sites:" line, printed to four significant digits), from the seed of
    the input."""
    for ln in out:
        if ln.startswith("sites:"):
            return [float(t.replace("D", "E")) for t in ln.split()[1:]]
    return None


def main():
    inp = open(os.path.join(REF, "be_vb1_jcas24_vmc_j_opt_lin.inp")).read()
    out = open(os.path.join(REF, "be_vb1_jcas24_vmc_j_opt_lin.out")).read().splitlines()
    runs = []
    # iteration 1 uses the input's own Jastrow
    params = None
    i = 0
    while i < len(out):
        ln = out[i]
        m = re.match(r"Beginning optimization iteration #\s*(\d+)", ln)
        if m:
            it = int(m.group(1))
            # find the next "Final results" and "total E =" line
            j = i
            while not out[j].startswith("Final results after"):
                j += 1
            passes = float(out[j].split()[3])
            while not out[j].startswith("total E ="):
                j += 1
            v = out[j].split()
            e, err, sig, sigerr = float(v[3]), float(v[5]), float(v[8]), float(v[10])
            runs.append(dict(iteration=it, params=params, energy=e, error=err,
                             sigma=sig, sigma_error=sigerr, passes=passes,
                             acceptance=_acceptance(out, j)))
        if ln.startswith("Wave function for next iteration #"):
            j = i + 1
            while out[j].strip() != "parameters":
                j += 1
            rows = []
            for k in range(3):
                nums = []
                for tok in out[j + 1 + k].split():
                    try:
                        nums.append(float(tok))
                    except ValueError:
                        break
                rows.append(nums)
            params = dict(a=rows[0], b=rows[1], c=rows[2])
        i += 1
    json.dump(dict(input=inp, runs=runs, sites=_sites(out)), open(os.path.join(HERE, "be_transcript.json"), "w"),
              indent=1)
    for r in runs:
        print(r["iteration"], r["energy"], r["error"], r["sigma"], r["passes"],
              None if r["params"] is None else len(r["params"]["c"]))


if __name__ == "__main__":
    main()
