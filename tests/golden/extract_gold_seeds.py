"""Write tests/golden/gold_rays.json from the reference's gold ray trace.

Input: test/tests/study/gold/openmc_study_rays.e of the reference (an
Exodus II / NetCDF-classic data file produced by the reference's
`study_run` regression, test/tests/study/tests:1-14).  It holds one EDGE2
record per element visit of the 30 rays of batch 3 with the ray aux data
registered in src/userobjects/OpenMCStudy.C:190-197.

Output: per record the ray id, the visit index (`intersections`) and the aux
data (energy, weight, id, n_progeny, particle type and the three seeds as
exact integers).  Only data is written; nothing of the reference is run.

Usage: python tests/golden/extract_gold_seeds.py /path/to/openmc_study_rays.e
"""
import json
import os
import sys

import numpy as np
from scipy.io import netcdf_file


def main(path):
    f = netcdf_file(path, "r", mmap=False)
    names = [b"".join(x).decode().strip() for x in f.variables["name_elem_var"][:]]
    cols = {n: np.array(f.variables[f"vals_elem_var{i + 1}eb1"][0]) for i, n in enumerate(names)}
    order = np.lexsort((cols["intersections"], cols["ray_id"]))
    rec = []
    for i in order:
        rec.append({
            "ray_id": int(cols["ray_id"][i]),
            "intersections": int(cols["intersections"][i]),
            "energy": float(cols["aux_energy"][i]),
            "weight": float(cols["aux_weight"][i]),
            "id": int(cols["aux_id"][i]),
            "n_progeny": int(cols["aux_n_progeny"][i]),
            "particle_type": int(cols["aux_particle_type"][i]),
            "seed_tracking": int(cols["aux_seed_tracking"][i]),
            "seed_source": int(cols["aux_seed_source"][i]),
            "seed_urr": int(cols["aux_seed_URR"][i]),
        })
    out = {
        "source": "test/tests/study/gold/openmc_study_rays.e",
        "settings": {"particles": 30, "batches": 3, "inactive": 1, "batch_in_file": 3,
                     "master_seed": 1, "stride": 152917},
        "n_records": len(rec),
        "records": rec,
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gold_rays.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, separators=(",", ":"))
    print(f"wrote {dst}: {len(rec)} records")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/tests/study/gold/openmc_study_rays.e")
