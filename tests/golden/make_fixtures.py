"""Writes the small data fixtures under tests/golden/ from PolyFields.jl's example seed
files (examples/seeds/*.json in the reference tree) so that no test reads the reference
at run time.  Run once in a checkout that has the reference at /root/reference:

    python tests/golden/make_fixtures.py

Each seed file stores ``fields[mid][z][y][x]`` (src/fileio.jl:107-133); the arrays are
saved as float64 grids of shape (Nz, Ny, Nx) together with grid, crystal and cell.
"""
import json
import os

import numpy as np

REF = "/root/reference/examples/seeds"
HERE = os.path.dirname(os.path.abspath(__file__))


def convert(name):
    with open(os.path.join(REF, name + ".json")) as fh:
        d = json.load(fh)
    out = {
        "grid": np.array(d["grid"], dtype=np.int64),
        "dim": np.array(d["dim"], dtype=np.int64),
        "cell_params": np.array(d["cell_params"], dtype=np.float64),
        "crystal": np.array(d["crystal"]),
        "mids": np.array(d["mids"], dtype=np.int64),
    }
    for mid in d["mids"]:
        out[f"field_{mid}"] = np.array(d["fields"][str(mid)], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, f"seed_{name}.npz"), **out)


if __name__ == "__main__":
    for n in ("lam1d", "cyl2d", "abc_blend_1d"):
        convert(n)
        print("wrote", n)
    # one seed file kept verbatim as a fixture of the JSON field-file format
    import shutil
    shutil.copyfile(os.path.join(REF, "lam1d.json"), os.path.join(HERE, "seed_lam1d.json"))
    print("copied lam1d.json")
