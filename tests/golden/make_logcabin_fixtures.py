"""Write the logcabin fixtures from the reference tree (run in a container
that has /root/reference; the GPU tests read only the files written here).

- logcabin.obj.gz, logcabin.mtl: the scene of runLogcabin.sh:3
  (objs/funyz/logcabin.{obj,mtl}), unchanged.
- logcabin_reference_image.npz: results/logcabin.bmp as data: `header` (the
  54 header bytes) and `bgr` (1024 x 1024 x 3 uint8, file row order, i.e.
  bottom-up rows, BGR).
- logcabin_case.json: the command-line flags of runLogcabin.sh:3 and the
  counts and bounds that results/logcabin.out:23-38 prints.
"""
import gzip
import json
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    obj = open(os.path.join(REF, "objs/funyz/logcabin.obj"), "rb").read()
    with gzip.GzipFile(os.path.join(OUT, "logcabin.obj.gz"), "wb", mtime=0) as f:
        f.write(obj)
    mtl = open(os.path.join(REF, "objs/funyz/logcabin.mtl"), "rb").read()
    open(os.path.join(OUT, "logcabin.mtl"), "wb").write(mtl)
    raw = open(os.path.join(REF, "results/logcabin.bmp"), "rb").read()
    header = np.frombuffer(raw[:54], np.uint8)
    bgr = np.frombuffer(raw[54:54 + 1024 * 1024 * 3], np.uint8).reshape(1024, 1024, 3)
    np.savez_compressed(os.path.join(OUT, "logcabin_reference_image.npz"), header=header, bgr=bgr)
    log = open(os.path.join(REF, "results/logcabin.out")).read()
    cmd = open(os.path.join(REF, "runLogcabin.sh")).read().splitlines()[2]
    args = cmd.split("cuda_bvh", 1)[1].split()
    counts = {k: int(re.search(r"num of %s : (\d+)" % k, log).group(1)) for k in ("vertices", "triangles", "materials")}
    m = re.search(r"found min\(([^,]+), ([^,]+) , ([^)]+)\)found max\(([^,]+), ([^,]+) , ([^)]+)\)", log)
    vals = [float(x) for x in m.groups()]
    faces = [int(x) for x in re.findall(r"Number of faces in mesh (\d+)", log)]
    case = {"args": args, "counts": counts, "normals": 4703, "shape_faces": faces,
            "bounds_min": vals[:3], "bounds_max": vals[3:], "spp": 20}
    json.dump(case, open(os.path.join(OUT, "logcabin_case.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
