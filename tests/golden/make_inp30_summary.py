"""Writes tests/golden/inp30_summary.json from the reference's only in-tree
mesh fixture, inputFiles/lidDrivenCavity/lidDrivenCavity_NE27000.inp (a 30^3
hex27 cavity with sinh clustering 2.0). The file itself stays in the
reference; the summary keeps what the tests compare: header values, counts,
the 31 axis positions (7 decimals as written by the reference generator),
SHA-256 digests of the connectivity and BC-face tables (int32, 0-based), the
zero-pressure node and the monitor point. Parsed as plain text; nothing from
the reference is executed."""
import hashlib
import json
import os
import sys

import numpy as np

SRC = "/root/reference/inputFiles/lidDrivenCavity/lidDrivenCavity_NE27000.inp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "inp30_summary.json")


def main(src=SRC, out=OUT):
    lines = open(src).read().splitlines()
    head = {}
    for ln in lines[2:20]:
        k, v = ln.split(":", 1)
        head[k.strip()] = float(v.split()[0])
    NE, NCN = int(head["NE"]), int(head["NCN"])
    c0 = 22
    coord = np.array([[float(t) for t in ln.split()[1:4]] for ln in lines[c0:c0 + NCN]])
    e0 = c0 + NCN + 2
    conn = np.array([[int(t) - 1 for t in ln.split()[1:9]] for ln in lines[e0:e0 + NE]], dtype=np.int32)
    i = e0 + NE + 2
    nbc = int(lines[i].split(":")[1]); i += 1
    bcs = []
    for k in range(nbc):
        vals = lines[i + k].split(":", 1)[1].replace(":", " ").split()
        bcs.append([float(v) for v in vals[1:4]])
    i += nbc + 1
    nvel = int(lines[i].split(":")[1]); i += 1
    nout = int(lines[i].split(":")[1]); i += 3
    faces = np.array([[int(t) - 1 for t in lines[i + k].split()[:3]] for k in range(nvel)], dtype=np.int32)
    i += nvel + 2 + nout + 2
    zero_p = int(lines[i]) - 1
    mon = [float(t) for t in lines[i + 3].split()]
    summary = dict(
        source="inputFiles/lidDrivenCavity/lidDrivenCavity_NE27000.inp", header=head, NE=NE, NCN=NCN,
        axis_x=sorted(set(coord[:, 0].tolist())), axis_y=sorted(set(coord[:, 1].tolist())),
        axis_z=sorted(set(coord[:, 2].tolist())), coord_sum=float(coord.sum()),
        conn_sha256=hashlib.sha256(conn.tobytes()).hexdigest(),
        faces_sha256=hashlib.sha256(faces.tobytes()).hexdigest(), n_vel_faces=nvel, n_out_faces=nout,
        BCstr=bcs, zeroP=zero_p, monitor=mon,
        faces_per_face_bc={f"{f}/{b}": int(((faces[:, 1] == f) & (faces[:, 2] == b)).sum())
                           for f in range(6) for b in range(nbc) if ((faces[:, 1] == f) & (faces[:, 2] == b)).any()},
    )
    with open(out, "w") as fh:
        json.dump(summary, fh, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main(*sys.argv[1:])
