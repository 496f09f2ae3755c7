"""Writes result_goldens.npz: the numeric fields of the reference's golden
Result files (test/data/NAMD/{tmao_tmao,water_tmao,water_water}.json,
test/data/PDB/tmao_tmao.json, test/data/Gromacs/{EMI_DCA,EMI_EMI}.json).

These files were written by the reference's save() (src/results.jl:533-538,
JSON3 layout).  They pin finalresults! (src/results.jl:311-469): every derived
field follows from the raw counts.  Run from the repository root:
    python tests/golden/make_result_goldens.py /root/reference
"""
import json
import os
import sys

import numpy as np

FILES = ["NAMD/tmao_tmao.json", "NAMD/water_tmao.json", "NAMD/water_water.json", "PDB/tmao_tmao.json",
         "Gromacs/EMI_DCA.json", "Gromacs/EMI_EMI.json"]
VEC = ["d", "md_count", "md_count_random", "coordination_number", "coordination_number_random", "mddf", "kb",
       "rdf_count", "rdf_count_random", "sum_rdf_count", "sum_rdf_count_random", "rdf", "kb_rdf"]
GROUPS = ["solute_group_count", "solvent_group_count", "solute_group_count_random", "solvent_group_count_random"]


def main(ref_root):
    out, meta = {}, {}
    for rel in FILES:
        key = rel.replace("/", "__").replace(".json", "")
        with open(os.path.join(ref_root, "test", "data", rel)) as fh:
            d = json.load(fh)
        for v in VEC:
            out[f"{key}.{v}"] = np.array(d[v], dtype=np.float64)
        for g in GROUPS:
            out[f"{key}.{g}"] = np.array(d[g], dtype=np.float64)
        out[f"{key}.volume_shell"] = np.array(d["volume"]["shell"], dtype=np.float64)
        f0 = d["files"][0]
        meta[key] = {
            "Version": d["Version"], "nbins": d["nbins"], "dbulk": d["dbulk"], "cutoff": d["cutoff"],
            "autocorrelation": d["autocorrelation"],
            "solute": {k: d["solute"][k] for k in ("nmols", "natomspermol", "custom_groups")},
            "solvent": {k: d["solvent"][k] for k in ("nmols", "natomspermol", "custom_groups")},
            "density": d["density"],
            "volume": {k: d["volume"][k] for k in ("total", "bulk", "domain")},
            "options": f0["options"], "irefatom": f0["irefatom"], "lastframe_read": f0["lastframe_read"],
            "nframes_read": f0["nframes_read"], "frame_weights": f0["frame_weights"], "weights": d["weights"],
        }
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "result_goldens.npz"), **out)
    with open(os.path.join(here, "result_goldens.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print(len(out), "arrays")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
