"""Extract known-answer vectors from the reference's own test inputs.

Run in the build container (reads the reference tree); the output
fixtures.json is committed and is the only thing the tests read.

Each restart file `.q_2d` of the reference's TESTS/ holds conservative
variables written by the reference's example generators for a physical state
that the accompanying IMEX_SfloW2D.inp deck and the survey of the generator
scripts determine: (h, alpha_s, T) -> q via qp_to_qc
(src/constitutive_2d.f90:1709-2140), written with format e20.12
(src/inpout_2d.f90:5814).  We record the first wet cell of each file.
"""
import json
import os

REF = "/root/reference"
CASES = [
    # file, deck parameters needed by qp_to_qc, physical state of the cell
    dict(file="TESTS/TEST_DAMBREAK/TEST_OFF_FINH/example_2D_0000.q_2d",
         n_solid=1, rho_s=[2500.0], sp_heat_s=[1617.0], sp_heat_a=998.0,
         sp_gas_const_a=287.051, pres=101300.0, energy_flag=True, alpha_flag=False,
         h=0.4, alphas=[0.58], T=300.0, n_cols_used=4 + 1),
    dict(file="TESTS/TEST_BOX/INIT_BOX_0000.q_2d",
         n_solid=1, rho_s=[2000.0], sp_heat_s=[1617.0], sp_heat_a=998.0,
         sp_gas_const_a=287.051, pres=101300.0, energy_flag=True, alpha_flag=False,
         h=200.0, alphas=[0.5], T=300.0, n_cols_used=4 + 1),
]


# Input data files of the reference's TESTS decks that the device-versus-oracle
# runs start from (restart states and their DEMs), kept gzip-compressed.
INPUT_FILES = [
    ("TESTS/TEST_BW/INIT_BW_0000.q_2d", "INIT_BW_0000.q_2d.gz"),
    ("TESTS/TEST_BW/topography_dem.asc", "TEST_BW_topography_dem.asc.gz"),
    ("TESTS/TEST_2D/INIT_2D_0000.q_2d", "INIT_2D_0000.q_2d.gz"),
    ("TESTS/TEST_2D/topography_dem.asc", "TEST_2D_topography_dem.asc.gz"),
]


def copy_inputs(here):
    import gzip
    for src, dst in INPUT_FILES:
        with open(os.path.join(REF, src), "rb") as f:
            data = f.read()
        with gzip.GzipFile(os.path.join(here, dst), "wb", compresslevel=9, mtime=0) as g:
            g.write(data)


def main():
    out = []
    for c in CASES:
        path = os.path.join(REF, c["file"])
        with open(path) as f:
            for line in f:
                vals = [float(v) for v in line.split()]
                if len(vals) >= 2 + c["n_cols_used"] and vals[2] > 0.0:
                    break
        rec = dict(c)
        rec["x"], rec["y"] = vals[0], vals[1]
        rec["q"] = vals[2:2 + c["n_cols_used"]]
        out.append(rec)
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))
    copy_inputs(here)


if __name__ == "__main__":
    main()
