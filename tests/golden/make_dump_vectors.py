"""Writes dump_vectors.json: how the reference's own binary reader
(scripts/binary_reader/vpic_binary_reader.py of lanl/vpic, read from
/root/reference in this container only) parses the small dumps of
dump_cases.py, with the sha256 of each file.  The format has no other
executable statement in the reference (its writers need the whole
simulation object), so this is the fixture that pins libvpichip's writers
to what the reference's readers expect.

Run from the repository root:  python tests/golden/make_dump_vectors.py
"""
import hashlib
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
sys.path.insert(0, HERE)
sys.path.insert(0, "/root/reference/scripts/binary_reader")

import numpy as np  # noqa: E402

import dump_cases  # noqa: E402
import vpic_binary_reader as R  # noqa: E402


def plain(v):
    if isinstance(v, (np.integer,)):
        return int(v)
    if isinstance(v, (np.floating,)):
        return float(v)
    if isinstance(v, (tuple, list)):
        return [plain(x) for x in v]
    return v


def main():
    out = {}
    with tempfile.TemporaryDirectory() as d:
        names = dump_cases.write_all(d)
        for kind in ("grid", "fields", "hydro", "particles"):
            fn = names[kind]
            with open(fn, "rb") as fh:
                hdr = R.read_header(fh)
                arr = R.read_array_header(fh, hdr)
                rec = {"header": {k: plain(v) for k, v in hdr.items()},
                       "array": {k: plain(v) for k, v in arr.items()}}
                if kind in ("fields", "hydro"):
                    data = R.read_data(fh, arr["dim"], hdr["floatfmt"],
                                       int(arr["no_fields"]))
                    rec["field_sums"] = [float(np.sum(x, dtype=np.float64))
                                         for x in data]
                if kind == "particles":
                    rec["particles"] = []
                    for _ in range(arr["dim"][0]):
                        rec["particles"].append(
                            plain(R.get_next_particle(fh, hdr)))
            with open(fn, "rb") as fh:
                rec["sha256"] = hashlib.sha256(fh.read()).hexdigest()
            out[kind] = rec
        with open(names["energies"], "rb") as fh:
            out["energies"] = {"sha256": hashlib.sha256(fh.read()).hexdigest()}
        snames = dump_cases.write_strided(d)
        for name in dump_cases.STRIDED:
            fn = snames[name]
            with open(fn, "rb") as fh:
                hdr = R.read_header(fh)
                arr = R.read_array_header(fh, hdr)
                data = R.read_data(fh, arr["dim"], hdr["floatfmt"],
                                   int(arr["no_fields"]))
                rec = {"header": {k: plain(v) for k, v in hdr.items()},
                       "array": {k: plain(v) for k, v in arr.items()},
                       "field_sums": [float(np.sum(x, dtype=np.float64))
                                      for x in data]}
            with open(fn, "rb") as fh:
                rec["sha256"] = hashlib.sha256(fh.read()).hexdigest()
            out["strided_" + name] = rec
        with open(snames["global"], "rb") as fh:
            out["global_header"] = {
                "sha256": hashlib.sha256(fh.read()).hexdigest()}
    with open(os.path.join(HERE, "dump_vectors.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "dump_vectors.json"))


if __name__ == "__main__":
    main()
