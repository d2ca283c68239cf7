"""Independent numpy cavity-mesh generator used by the tests.

It follows the reference cube generator
oldFiles/meshGenerators&Converters/HexaMeshGeneratorInACube_GeneratesCornerNodes.m
(GEN): corner nodes x-fastest (GEN:134-148), elements x-fastest with the
corner order n1..n8 (GEN:152-169), sinh clustering (GEN:100-131), velocity BC
faces bottom/left/right/front/back/top with the lid (BC 2) last (GEN:218-285),
zero-pressure node (NEy/2)*NxC + (NxC+1)/2 (1-based, GEN:294) and monitor
point (0.5, 0.5, 0.5). All indices returned here are 0-based.

The product library has its own C++ generator; the tests check that both give
identical arrays, and that this one reproduces the reference's shipped mesh
inputFiles/lidDrivenCavity/lidDrivenCavity_NE27000.inp.
"""
import numpy as np


def axis(n_el, L, cluster):
    nc = n_el + 1
    if cluster != 0.0 and n_el % 2:
        raise ValueError("sinh clustering needs an even element count (GEN:41)")
    if cluster == 0.0:
        return np.array([(i * L) / (nc - 1) for i in range(nc)], dtype=np.float64)
    mx = np.sinh(cluster)
    d = np.zeros(nc)
    half = (nc + 1) // 2
    for i in range(1, half + 1):
        xx = (1.0 * (i - 1)) / (half - 1)
        d[i - 1] = L / 2.0 / mx * np.sinh(cluster * xx)
    for i in range(half, nc + 1):
        d[i - 1] = L - d[nc + 1 - i - 1]
    return d


def cavity(N, cluster=0.0, lid=(1.0, 0.0, 0.0)):
    nx = ny = nz = N
    ncx, ncy, ncz = nx + 1, ny + 1, nz + 1
    dx, dy, dz = axis(nx, 1.0, cluster), axis(ny, 1.0, cluster), axis(nz, 1.0, cluster)
    kk, jj, ii = np.meshgrid(np.arange(ncz), np.arange(ncy), np.arange(ncx), indexing="ij")
    coord = np.stack([dx[ii.ravel()], dy[jj.ravel()], dz[kk.ravel()]], axis=1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    n1 = (k * ncx * ncy + j * ncx + i).ravel()
    n2 = n1 + 1
    n3 = n1 + nx + 2
    n4 = n3 - 1
    lay = ncx * ncy
    conn = np.stack([n1, n2, n3, n4, n1 + lay, n2 + lay, n3 + lay, n4 + lay], axis=1).astype(np.int32)
    faces = []
    e = lambda i_, j_, k_: k_ * nx * ny + j_ * nx + i_
    for jy in range(ny):                      # bottom, face 1 (0-based 0)
        for ix in range(nx):
            faces.append((e(ix, jy, 0), 0, 0))
    for kz in range(nz):                      # left, face 5
        for jy in range(ny):
            faces.append((e(0, jy, kz), 4, 0))
    for kz in range(nz):                      # right, face 3
        for jy in range(ny):
            faces.append((e(nx - 1, jy, kz), 2, 0))
    for kz in range(nz):                      # front, face 2
        for ix in range(nx):
            faces.append((e(ix, 0, kz), 1, 0))
    for kz in range(nz):                      # back, face 4
        for ix in range(nx):
            faces.append((e(ix, ny - 1, kz), 3, 0))
    for jy in range(ny):                      # top (lid), face 6, BC 2
        for ix in range(nx):
            faces.append((e(ix, jy, nz - 1), 5, 1))
    zero_p = (ny // 2) * ncx + (ncx + 1) // 2 - 1
    return dict(coord=coord, conn=conn, faces=np.array(faces, dtype=np.int32),
                BCstr=np.array([[0.0, 0.0, 0.0], list(lid)], dtype=np.float64),
                zeroP=zero_p, mon=np.array([0.5, 0.5, 0.5]))


def write_inp(path, mesh, NGP=8, dt=0.01, t_final=1.0, maxIter=3, tolerance=1e-3, converge=1e-6,
              density=1.0, viscosity=0.01, coord_fmt="%18.7f"):
    """Write the .inp layout read by BCH:528-725 (with the `converge` line, INP:15)."""
    coord, conn, faces = mesh["coord"], mesh["conn"], mesh["faces"]
    L = ["3D Lid-driven cavity problem ", "=" * 48,
         f"eType    : 1 ", f"NE       : {conn.shape[0]} ", f"NCN      : {coord.shape[0]} ",
         "NENv     : 27 ", "NENp     : 8 ", f"NGP      : {NGP} ", "alpha    : 1.000000",
         f"dt       : {dt:f}", "t_ini    : 0.000000 ", f"t_final  : {t_final:f} ",
         f"maxIter  : {maxIter} ", f"tolerance: {tolerance:f}", f"converge : {converge:f} ",
         "isRestart: 0", f"density  : {density:f} ", f"viscosity: {viscosity:f} ", "fx       : 0.0 ",
         "fy       : 0.0 ", "=" * 48, "Corner Node No         x                y                z"]
    for n, (x, y, z) in enumerate(coord):
        L.append(("%9i " + coord_fmt + " " + coord_fmt + " " + coord_fmt) % (n + 1, x, y, z))
    L += ["=" * 48, "Elem No   corner1  corner2  corner3  corner4  corner5  corner6  corner7  corner8"]
    for e, c in enumerate(conn):
        L.append("%6i  " % (e + 1) + " ".join("%8i" % (v + 1) for v in c) + " ")
    L += ["=" * 48, "BCs (Number of specified BCs, their types and strings) ",
          f"nBC       : {len(mesh['BCstr'])} "]
    for b, s in enumerate(mesh["BCstr"]):
        L.append(f"BC {b + 1}      : 1  {s[0]:.1f} : {s[1]:.1f} : {s[2]:.1f}")
    L += ["=" * 48, f"nVelFaces : {faces.shape[0]} ", "nOutFaces : 0 ", "=" * 48,
          "Velocity BC (Elem# Face# BC#)"]
    for (e, f, b) in faces:
        L.append("%5d    %d    %d" % (e + 1, f + 1, b + 1))
    L += ["=" * 48, "Outflow BC (Elem# Face# BC#)", "=" * 48,
          "Node number where pressure is taken to be zero", "%d" % (mesh["zeroP"] + 1), "=" * 48,
          "Monitor point coordinates", "%g  %g  %g" % tuple(mesh["mon"])]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
