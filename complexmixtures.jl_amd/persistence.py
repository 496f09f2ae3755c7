"""Result persistence in the reference's JSON layout.

save/load mirror src/results.jl:533-588 (JSON3 with StructTypes.Struct: the
fields of Result in declaration order, compact, "Version" first) and the
version gate of src/io.jl:31-87.  Files written here load in the reference
and the reference's files load here; files from <= 2.17.1 (no random group
counts, src/legacy/results-2.17.1.jl) load with empty random group counts.
"""
import json
import math
import re
import warnings

import numpy as np

from .options import ArgumentError, Options, TrajectoryFileOptions
from .results import VERSION, Density, Result, Volume
from .selection import AtomSelection


def _parse_version(v):
    m = re.match(r"^(\d+)\.(\d+)\.(\d+)(?:-([0-9A-Za-z.\-]+))?", v)
    if not m:
        return (1, 0, 0, 1, ())
    pre = m.group(4)
    # a pre-release sorts before the release (Julia VersionNumber ordering)
    return (int(m.group(1)), int(m.group(2)), int(m.group(3)), 0 if pre else 1, tuple((pre or "").split(".")))


def _num(x):
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    r = repr(x)
    return r


def _vec(a):
    return "[" + ",".join(_num(x) for x in np.asarray(a, dtype=np.float64).ravel()) + "]"


def _mat(a):
    return "[" + ",".join(_vec(row) for row in np.asarray(a)) + "]"


def _js(x):
    return json.dumps(x, separators=(",", ":"))


def _selection(s):
    return ("{" + f'"nmols":{s.nmols},"natomspermol":{s.natomspermol},"indices":{_js(list(s.indices))},'
            f'"custom_groups":{_js(s.custom_groups)},"group_atom_indices":{_js(s.group_atom_indices)},'
            f'"group_names":{_js(list(s.group_names))}' + "}")


def _options(o):
    parts = []
    for k, v in o.to_dict().items():
        parts.append(f'"{k}":' + (_num(v) if isinstance(v, float) else _js(v)))
    return "{" + ",".join(parts) + "}"


def dumps(R):
    f = ",".join("{" + f'"filename":{_js(x.filename)},"options":{_options(x.options)},"irefatom":{x.irefatom},'
                 f'"lastframe_read":{x.lastframe_read},"nframes_read":{x.nframes_read},'
                 f'"frame_weights":{_vec(x.frame_weights)}' + "}" for x in R.files)
    parts = [
        f'"Version":{_js(str(R.Version))}', f'"nbins":{R.nbins}', f'"dbulk":{_num(R.dbulk)}',
        f'"cutoff":{_num(R.cutoff)}', f'"d":{_vec(R.d)}', f'"md_count":{_vec(R.md_count)}',
        f'"md_count_random":{_vec(R.md_count_random)}', f'"coordination_number":{_vec(R.coordination_number)}',
        f'"coordination_number_random":{_vec(R.coordination_number_random)}', f'"mddf":{_vec(R.mddf)}',
        f'"kb":{_vec(R.kb)}', f'"autocorrelation":{_js(bool(R.autocorrelation))}',
        f'"solute":{_selection(R.solute)}', f'"solvent":{_selection(R.solvent)}',
        f'"solute_group_count":{_mat(R.solute_group_count)}', f'"solvent_group_count":{_mat(R.solvent_group_count)}',
        f'"solute_group_count_random":{_mat(R.solute_group_count_random)}',
        f'"solvent_group_count_random":{_mat(R.solvent_group_count_random)}',
        f'"rdf_count":{_vec(R.rdf_count)}', f'"rdf_count_random":{_vec(R.rdf_count_random)}',
        f'"sum_rdf_count":{_vec(R.sum_rdf_count)}', f'"sum_rdf_count_random":{_vec(R.sum_rdf_count_random)}',
        f'"rdf":{_vec(R.rdf)}', f'"kb_rdf":{_vec(R.kb_rdf)}',
        '"density":{' + f'"solute":{_num(R.density.solute)},"solvent":{_num(R.density.solvent)},'
                        f'"solvent_bulk":{_num(R.density.solvent_bulk)}' + "}",
        '"volume":{' + f'"total":{_num(R.volume.total)},"bulk":{_num(R.volume.bulk)},'
                       f'"domain":{_num(R.volume.domain)},"shell":{_vec(R.volume.shell)}' + "}",
        f'"files":[{f}]', f'"weights":{_vec(R.weights)}',
    ]
    return "{" + ",".join(parts) + "}"


def save(filename, R):
    """save(filename, R), src/results.jl:533-541 (save(R, filename) also accepted)."""
    if isinstance(filename, Result):
        filename, R = R, filename
    with open(filename, "w") as fh:
        fh.write(dumps(R))
    return f"Results saved in JSON file: {filename}"


def _get_version(filename):
    with open(filename) as fh:
        head = fh.read(4096).split(",", 1)[0]
    m = re.search(r'"Version":"([^"]*)"', head)
    return m.group(1) if m else "1.0.0"


def _check_version(filename):
    """src/io.jl:31-87"""
    v = _get_version(filename)
    pv, cur = _parse_version(v), _parse_version(VERSION)
    if pv > cur:
        raise ArgumentError("\n\nTrying to load a json result file created with a newer version of ComplexMixtures. "
                            f"\nThis can cause unpredictable errors. \n\nCurrent version of ComplexMixtures: {VERSION}\n"
                            f"Version used to create the output .json file: {v}\n\nPlease update ComplexMixtures and "
                            "try again.\n")
    if pv < _parse_version("2.0.0"):
        raise ArgumentError("\n\nTrying to load a json result created with an older, and incompatible, version of "
                            f"ComplexMixtures.\n\nCurrent version of ComplexMixtures: {VERSION}\n"
                            f"Version used to create the output .json file: {v}\n")
    legacy = _parse_version("2.0.0") < pv <= _parse_version("2.17.1")
    if legacy:
        warnings.warn(f"\nReading json result created with {v}. \n\nThis version didn't support KBI contribution "
                      "decomposition, introduced in v2.18.0. \n\nThe `solute_group_count_random` and "
                      "`solvent_group_count_random` arrays will be empty.\n")
    return legacy


def _parse_constant(c):
    return {"NaN": math.nan, "Inf": math.inf, "-Inf": -math.inf, "Infinity": math.inf,
            "-Infinity": -math.inf}[c]


def load(filename):
    """load(filename), src/results.jl:567-588"""
    legacy = _check_version(filename)
    try:
        with open(filename) as fh:
            txt = fh.read()
        txt = re.sub(r"(?<=[:,\[])\s*(-?)Inf\b", lambda m: f"{m.group(1)}Infinity", txt)
        d = json.loads(txt, parse_constant=_parse_constant)
        sel = lambda x: AtomSelection.from_dict(x)
        su, sv = sel(d["solute"]), sel(d["solvent"])
        files = [TrajectoryFileOptions(filename=x["filename"], options=Options.from_dict(x["options"]),
                                       irefatom=int(x["irefatom"]), lastframe_read=int(x["lastframe_read"]),
                                       nframes_read=int(x["nframes_read"]),
                                       frame_weights=[float(w) for w in x["frame_weights"]]) for x in d["files"]]
        gu = np.array(d["solute_group_count"], dtype=np.float64)
        gv = np.array(d["solvent_group_count"], dtype=np.float64)
        R = Result(nbins=d["nbins"], dbulk=d["dbulk"], cutoff=d["cutoff"], autocorrelation=d["autocorrelation"],
                   solute=su, solvent=sv, n_groups_solute=len(gu), n_groups_solvent=len(gv), files=files,
                   weights=[float(w) for w in d["weights"]])
        R.Version = d["Version"]
        for k in Result._vec_fields:
            getattr(R, k)[:] = np.array(d[k], dtype=np.float64)
        R.solute_group_count[:] = gu
        R.solvent_group_count[:] = gv
        if not legacy:
            R.solute_group_count_random[:] = np.array(d["solute_group_count_random"], dtype=np.float64)
            R.solvent_group_count_random[:] = np.array(d["solvent_group_count_random"], dtype=np.float64)
        else:
            R.solute_group_count_random = np.zeros((0, R.nbins))
            R.solvent_group_count_random = np.zeros((0, R.nbins))
        R.density = Density(**{k: float(v) for k, v in d["density"].items()})
        v = d["volume"]
        R.volume = Volume(float(v["total"]), float(v["bulk"]), float(v["domain"]),
                          np.array(v["shell"], dtype=np.float64))
    except (KeyError, TypeError, ValueError, json.JSONDecodeError) as e:
        raise ArgumentError(f"\n\nThe file {filename} does not appear to contain a Result object.\n") from e
    return R


def results_equal(R1, R2, rtol=0.0):
    """Field-wise comparison ignoring files, like src/compare.jl:26-52."""
    for k in Result._vec_fields + Result._group_fields:
        a, b = np.asarray(getattr(R1, k)), np.asarray(getattr(R2, k))
        if a.shape != b.shape:
            return False
        if rtol == 0.0:
            if not np.array_equal(a, b, equal_nan=True):
                return False
        elif not np.allclose(a, b, rtol=rtol, atol=0.0, equal_nan=True):
            return False
    for k in ("nbins", "dbulk", "cutoff", "autocorrelation"):
        if getattr(R1, k) != getattr(R2, k):
            return False
    if R1.solute != R2.solute or R1.solvent != R2.solvent:
        return False
    for part in ("density", "volume"):
        a, b = getattr(R1, part), getattr(R2, part)
        for k in vars(a):
            x, y = getattr(a, k), getattr(b, k)
            if not np.allclose(x, y, rtol=rtol, atol=0.0, equal_nan=True):
                return False
    return list(R1.weights) == list(R2.weights)
