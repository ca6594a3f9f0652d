"""Copies the input files of the reference's functional tests for three molecules into
tests/golden/<molecule>/ and extracts their statistical reference results
(tests/functional/statistical_references/<molecule>/<hamiltonian>/<trial>/<walkers>/results.h5:
the run parameters from `input_file`, mean / error of Energy, its components and Overlap, and of
the mixed and back-propagated OneRDM and DiagonalTwoRDM) with h5dump:
  BH  19 orbitals, 78 Cholesky vectors, RHF / UHF / CASCI-PHMSD / CASCI-NOMSD trials
  N2  28 orbitals, 134 Cholesky vectors, a 50-determinant CAS trial (PHMSD)
  Li  14 orbitals, 70 Cholesky vectors, a fully polarized ROHF trial (3 + 0 electrons)
Run where the reference's sources are; GPU runs only read the copies.

    python tests/golden/make_molecule_fixtures.py
"""
import json
import os
import re
import shutil
import subprocess

REF = "/root/reference/tests/functional"
H5DUMP = "/opt/conda/bin/h5dump"
HERE = os.path.dirname(os.path.abspath(__file__))
MOLECULES = {
    "BH": {
        "inputs": ["afqmc_H_rhf_closed.h5", "afqmc_H_rhf_collinear.h5", "afqmc_rhf_nomsd.h5",
                   "afqmc_uhf_nomsd.h5", "afqmc_casci_rhf_phmsd.h5", "afqmc_casci_uhf_phmsd.h5",
                   "afqmc_casci_rhf_nomsd.h5", "afqmc_casci_uhf_nomsd.h5"],
        "cases": [("dense_rhf_basis_closed", "rhf_nomsd", "closed"),
                  ("dense_rhf_basis_closed", "rhf_nomsd", "collinear"),
                  ("dense_rhf_basis_closed", "uhf_nomsd", "collinear"),
                  ("dense_rhf_basis_closed", "rcasci_rhf_phmsd", "collinear"),
                  ("dense_rhf_basis_closed", "rcasci_uhf_phmsd", "collinear"),
                  ("dense_rhf_basis_closed", "rcasci_rhf_nomsd", "closed"),
                  ("dense_rhf_basis_closed", "rcasci_rhf_nomsd", "collinear"),
                  ("dense_rhf_basis_closed", "rcasci_uhf_nomsd", "collinear")],
    },
    "N2": {
        "inputs": ["cas_basis_hamil.h5", "cas_wfn.h5"],
        "cases": [("cas_basis_hamil_closed", "cas_wfn_collinear", "collinear")],
    },
    "Li": {
        "inputs": ["hamil_closed.h5", "rohf_nomsd_polarized.h5"],
        "cases": [("hamil_closed", "rohf_wfn_polarized", "collinear")],
    },
}


def dump_value(path, dataset):
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", dataset, path], check=True,
                         capture_output=True, text=True).stdout
    body = out[out.index("DATA {"):]
    nums = re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", body.split("}")[0] + "}" + body.split("}")[1])
    return [float(x) for x in nums[1:]]   # drop the "(0)" index


def dump_string(path, dataset):
    out = subprocess.run([H5DUMP, "-d", dataset, path], check=True, capture_output=True,
                         text=True).stdout
    m = re.search(r'\(0\): "(.*)"\s*\}', out, re.S)
    return m.group(1).replace('\\"', '"')


def dump_array(path, dataset):
    """A complex dataset ({r, i} compound) as a nested list [[re, im], ...] with its shape."""
    out = subprocess.run([H5DUMP, "-y", "-w", "0", "-m", "%.17g", "-d", dataset, path], check=True,
                         capture_output=True, text=True).stdout
    shape = [int(x) for x in re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(([^)]*)\)", out).group(1).split(",")]
    body = out[out.index("DATA {") + len("DATA {"):]
    nums = [float(x) for x in re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", body)]
    n = 1
    for d in shape:
        n *= d
    assert len(nums) == 2 * n, (dataset, len(nums), shape)
    return {"shape": shape, "re": nums[0::2], "im": nums[1::2]}


def extract(molecule, spec):
    out = os.path.join(HERE, molecule.lower())
    os.makedirs(out, exist_ok=True)
    for f in spec["inputs"]:
        shutil.copyfile(os.path.join(REF, "afqmc_inputs", molecule, f), os.path.join(out, f))
    refs = {}
    for basis, trial, walkers in spec["cases"]:
        path = os.path.join(REF, "statistical_references", molecule, basis, trial, walkers,
                            "results.h5")
        if not os.path.exists(path):
            continue
        inp = json.loads(dump_string(path, "/input_file"))
        ex = inp["execute"]
        nranks = int(dump_value(path, "/num_ranks")[0])
        obs = {}
        for name in ("Energy", "OnebodyEnergy", "CoulombEnergy", "ExchangeEnergy", "Overlap"):
            mean = dump_value(path, f"/measurements/Stage0/{name}/mean")
            err = dump_value(path, f"/measurements/Stage0/{name}/error")
            obs[name] = {"mean": mean[:2], "error": err[:2]}
        # density-matrix observables of the mixed and back-propagation estimators
        rdm = {}
        est = ex.get("estimators") or {}
        groups = []
        if "mixed" in est:
            groups.append(("mixed", "MixedEstimator"))
        if "backprop" in est:
            for L in est["backprop"]["propagation_steps"]:
                groups.append((f"backprop/{L}", f"BackPropEstimator/Steps={L}"))
        for key, grp in groups:
            for o in ("OneRDM", "DiagonalTwoRDM"):
                try:
                    rdm.setdefault(key, {})[o] = {
                        "mean": dump_array(path, f"/measurements/Stage0/{grp}/{o}/mean"),
                        "error": dump_array(path, f"/measurements/Stage0/{grp}/{o}/error")}
                except subprocess.CalledProcessError:
                    pass
        refs[f"{trial}/{walkers}"] = {
            "rdm": rdm,
            "estimators": ex.get("estimators"),
            "hamiltonian": os.path.basename(ex["hamiltonian"]["filename"]),
            "wavefunction": os.path.basename(ex["wavefunction"]["filename"]),
            "walker_type": ex["walker_set"]["walker_type"],
            "timestep": ex["timestep"], "steps": ex["steps"],
            "walkers": ex.get("n_walkers_per_mpi_task", ex.get("num_walkers", 0)) * nranks,
            "num_ranks": nranks,
            "equilibration_steps": ex.get("equilibration_steps"),
            "measure_interval": ex.get("measure_interval"),
            "population_control_interval": ex.get("population_control_interval"),
            "walker_ortho_interval": ex.get("walker_ortho_interval"),
            "seed": inp.get("seed"),
            "observables": obs,
        }
    with open(os.path.join(out, "statistical_refs.json"), "w") as f:
        json.dump(refs, f, sort_keys=True)
    return refs


def main():
    for molecule, spec in MOLECULES.items():
        refs = extract(molecule, spec)
        for k, v in refs.items():
            print(molecule, k, v["walkers"], v["steps"], {o: d["mean"] for o, d in v["observables"].items()})


if __name__ == "__main__":
    main()
