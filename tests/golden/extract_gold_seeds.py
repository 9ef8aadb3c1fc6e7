"""Write tests/golden/gold_rays.json from the reference's gold ray trace.

Input: test/tests/study/gold/openmc_study_rays.e of the reference (an
Exodus II / NetCDF-classic data file produced by the reference's
`study_run` regression, test/tests/study/tests:1-14).  It holds one EDGE2
record per element visit of the 30 rays of batch 3 with the ray aux data
registered in src/userobjects/OpenMCStudy.C:190-197.

Output: per record the ray id, the visit index (`intersections`) and the aux
data (energy, weight, id, n_progeny, particle type and the three seeds as
exact integers) in gold_rays.json, and the file's layout (dimensions,
variables, element-variable names, attributes) in gold_rays_layout.json.
Only data is written; nothing of the reference is run.

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
    # the file's layout (dimensions, variables, element-variable names,
    # global attributes) for the trace writer's format test
    fixed = {k: v for k, v in f.dimensions.items() if v is not None}
    layout = {
        "dimensions_fixed": {k: v for k, v in fixed.items()
                             if k in ("len_string", "len_line", "four", "len_name", "num_dim", "num_el_blk",
                                      "num_nod_per_el1")},
        "dimensions_sized": {"num_nodes": "num_elem + number of rays", "num_elem": "visits",
                             "num_el_in_blk1": "visits", "num_elem_var": len(names)},
        "unlimited": [k for k, v in f.dimensions.items() if v is None],
        "variables": {k: [list(v.dimensions), v.typecode()] for k, v in f.variables.items()},
        "variable_attributes": {k: {a: (b.decode() if isinstance(b, bytes) else b) for a, b in v._attributes.items()}
                                for k, v in f.variables.items() if v._attributes},
        "time_whole": float(f.variables["time_whole"][0]),
        "connect1_elem_type": f.variables["connect1"].elem_type.decode(),
        "elem_var_names": names,
        "attributes": {k: (v.decode() if isinstance(v, bytes) else float(v) if isinstance(v, np.floating)
                           else int(v)) for k, v in f._attributes.items() if k != "title"},
        "counts": {"num_nodes": f.dimensions["num_nodes"], "num_elem": f.dimensions["num_elem"],
                   "rays": int(len(set(cols["ray_id"].tolist())))},
    }
    dst2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gold_rays_layout.json")
    with open(dst2, "w") as fh:
        json.dump(layout, fh, indent=1)
    print(f"wrote {dst2}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/tests/study/gold/openmc_study_rays.e")
