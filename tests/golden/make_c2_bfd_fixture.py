"""Extract the pseudopotential C2 wavefunction and VMC results of CHAMP's
shipped transcript (examples/c2_bfd_vtz_rhf_vmc.{inp,out}: old-format
input, numerical radial basis basis.1, Gaussian pseudopotential pseudopot1,
6-point quadrature, isc = 4 Jastrow) into tests/golden/c2_bfd_transcript.json
and tests/golden/c2_bfd_basis.1.gz.

Reads the reference's example files as text (data); nothing is executed.
The first "sites:" line (the starting configuration from the seed) is kept.
Each record: the Jastrow parameters of one VMC run (iteration 1: the input's
own; iteration k > 1: the "New wave function" block printed after iteration
k - 1), that run's "total E" line (energy, error, sigma, sigma error), its
potential and interaction energies, the number of passes and the acceptance
ratio.  The basis file (2000 grid points of 9 radial functions) is stored
gzip-compressed as it is.
"""
import gzip
import json
import os
import re

REF = "/root/reference/examples"
HERE = os.path.dirname(os.path.abspath(__file__))
STEM = "c2_bfd_vtz_rhf_vmc"


def _nums(line):
    out = []
    for tok in line.split():
        try:
            out.append(float(tok))
        except ValueError:
            break
    return out


def _after(out, j, prefix):
    while not out[j].startswith(prefix):
        j += 1
    return j


def _sites(out):
    """The first configuration CHAMP generated with its sites routine (the
    "sites:" line, printed to four significant digits), from the seed of
    the input."""
    for ln in out:
        if ln.startswith("sites:"):
            return [float(t.replace("D", "E")) for t in ln.split()[1:]]
    return None


def main():
    inp = open(os.path.join(REF, STEM + ".inp")).read()
    out = open(os.path.join(REF, STEM + ".out")).read().splitlines()
    runs, params = [], None
    for i, ln in enumerate(out):
        m = re.match(r"Beginning optimization iteration #\s*(\d+)", ln)
        if m:
            it = int(m.group(1))
            j = _after(out, i, "Final results after")
            passes = float(out[j].split()[3])
            j = _after(out, j, "total E =")
            v = out[j].split()
            jp = _after(out, j, "potential E =")
            ji = _after(out, j, "interaction E =")
            ja = _after(out, j, "acceptance")
            runs.append(dict(iteration=it, params=params, energy=float(v[3]), error=float(v[5]),
                             sigma=float(v[8]), sigma_error=float(v[10]), passes=passes,
                             potential=float(out[jp].split()[3]),
                             potential_error=float(out[jp].split()[5]),
                             interaction=float(out[ji].split()[3]),
                             interaction_error=float(out[ji].split()[5]),
                             acceptance=float(out[ja].split()[1])))
        if ln.startswith("New wave function:"):
            j = _after(out, i + 1, "Jastrow parameters:")
            rows = [_nums(out[j + 1 + k]) for k in range(3)]
            params = dict(a=rows[0], b=rows[1], c=rows[2])
    psp = open(os.path.join(REF, "pseudopot1")).read()
    json.dump(dict(input=inp, pseudopot1=psp, runs=runs, sites=_sites(out)),
              open(os.path.join(HERE, "c2_bfd_transcript.json"), "w"), indent=1)
    with open(os.path.join(REF, "basis.1"), "rb") as f, \
            gzip.GzipFile(os.path.join(HERE, "c2_bfd_basis.1.gz"), "wb", mtime=0) as g:
        g.write(f.read())
    for r in runs:
        print(r["iteration"], r["energy"], r["error"], r["sigma"], r["passes"], r["acceptance"],
              None if r["params"] is None else [len(r["params"][k]) for k in "abc"])


if __name__ == "__main__":
    main()
