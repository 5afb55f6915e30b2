"""Writes tests/golden/smoke_subset/: every 8th splat of the reference's real first-smoke scene
(smoke-assets/scaniverse-first-smoke, 94,406 Scaniverse splats with SAM3 roughness / metalness),
in the same manifest + little-endian sidecar format (schema scaniverse_first_smoke_splat_v1).

The rows are copied byte for byte; the manifest keeps the source's declared bounds and decode
notes, with the counts and byte lengths of the subset and ids renumbered 0..n-1 (the decoder
requires every original ID to be below the row count).  Run from the repository root with the
reference checked out at /root/reference:

    python tests/golden/make_smoke_subset.py
"""
import json
import os

import numpy as np

SRC = "/root/reference/smoke-assets/scaniverse-first-smoke"
NAME = "scaniverse-first-smoke"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "smoke_subset")
STEP = 8


def main() -> None:
    with open(os.path.join(SRC, NAME + ".json")) as f:
        man = json.load(f)
    n = man["splat_count"]
    keep = np.arange(0, n, STEP)
    m = int(keep.size)
    os.makedirs(DST, exist_ok=True)

    def take(path, comps, dtype="<f4"):
        a = np.fromfile(os.path.join(SRC, path), dtype=dtype).reshape(n, comps)
        return np.ascontiguousarray(a[keep])

    rows = take(man["payload"]["path"], 8)
    scales = take(man["shape"]["scales_path"], 3)
    rots = take(man["shape"]["rotations_path"], 4)
    rough = take(man["materials"]["roughness_path"], 1)
    metal = take(man["materials"]["metalness_path"], 1)
    ids = np.arange(m, dtype="<u4")
    files = {man["payload"]["path"]: rows, man["shape"]["scales_path"]: scales, man["shape"]["rotations_path"]: rots,
             man["materials"]["roughness_path"]: rough, man["materials"]["metalness_path"]: metal,
             man["ids"]["path"]: ids}
    for path, arr in files.items():
        arr.tofile(os.path.join(DST, path))
    man["splat_count"] = m
    man["payload"]["byte_length"] = int(rows.nbytes)
    man["ids"]["byte_length"] = int(ids.nbytes)
    man["ids"]["first_id"] = 0
    man["ids"]["last_id"] = m - 1
    man["shape"]["scales_byte_length"] = int(scales.nbytes)
    man["shape"]["rotations_byte_length"] = int(rots.nbytes)
    man["materials"]["roughness_byte_length"] = int(rough.nbytes)
    man["materials"]["metalness_byte_length"] = int(metal.nbytes)
    man["subset"] = {"of": NAME, "source_rows": n, "row_step": STEP,
                     "note": "every 8th row of the source, ids renumbered 0..n-1"}
    with open(os.path.join(DST, NAME + ".json"), "w") as f:
        json.dump(man, f, indent=2)
        f.write("\n")
    print(f"wrote {m} splats to {DST}")


if __name__ == "__main__":
    main()
