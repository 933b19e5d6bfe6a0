"""Small OBJ/MTL files written by the tests (a cube with two materials, uvs
and normals; a pentagon; a smooth UV sphere; files with faults)."""
from pathlib import Path

CUBE_MTL = """# two materials
newmtl red
Kd 0.8 0.1 0.2
Pr 0.3
newmtl grey
Kd 0.5 0.5 0.5
"""

# unit-half cube [-1,1]^3: 6 quads; +z face first with uvs (0,0) (1,0) (1,1) (0,1)
CUBE_OBJ = """mtllib cube.mtl
o cube
v -1 -1  1
v  1 -1  1
v  1  1  1
v -1  1  1
v -1 -1 -1
v  1 -1 -1
v  1  1 -1
v -1  1 -1
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 1
vn 0 0 -1
vn 1 0 0
vn -1 0 0
vn 0 1 0
vn 0 -1 0
usemtl red
f 1/1/1 2/2/1 3/3/1 4/4/1
f 6/1/2 5/2/2 8/3/2 7/4/2
usemtl grey
f 2/1/3 6/2/3 7/3/3 3/4/3
f 5/1/4 1/2/4 4/3/4 8/4/4
f 4/1/5 3/2/5 7/3/5 8/4/5
f -4/1/6 -3/2/6 -7/3/6 -8/4/6
"""

PENTAGON_OBJ = """v 0 0 0
v 1 0 0
v 1.3 1 0
v 0.5 1.6 0
v -0.3 1 0
f 1 2 3 4 5
"""


def write_cube(d: Path) -> Path:
    (d / "cube.mtl").write_text(CUBE_MTL)
    p = d / "cube.obj"
    p.write_text(CUBE_OBJ)
    return p


def uv_sphere_obj(n_lat: int = 12, n_lon: int = 24, r: float = 1.0) -> str:
    """A smooth UV sphere (y up): vertex (i, j) at latitude theta_i = pi i / n_lat
    and longitude phi_j = 2 pi j / n_lon, shared by the faces around it, with
    texture (u, v) = (j / n_lon, i / n_lat) and the unit position as normal.
    The seam column j = n_lon repeats j = 0's positions with u = 1."""
    import math
    lines = []
    for i in range(n_lat + 1):
        th = math.pi * i / n_lat
        for j in range(n_lon + 1):
            ph = 2 * math.pi * j / n_lon
            x, y, z = math.cos(ph) * math.sin(th), math.cos(th), math.sin(ph) * math.sin(th)
            lines.append(f"v {r * x:.9f} {r * y:.9f} {r * z:.9f}")
            lines.append(f"vn {x:.9f} {y:.9f} {z:.9f}")
            lines.append(f"vt {j / n_lon:.9f} {i / n_lat:.9f}")
    w = n_lon + 1
    for i in range(n_lat):
        for j in range(n_lon):
            a, b, c, d = i * w + j + 1, i * w + j + 2, (i + 1) * w + j + 2, (i + 1) * w + j + 1
            # counter-clockwise seen from outside
            lines.append(f"f {a}/{a}/{a} {d}/{d}/{d} {c}/{c}/{c} {b}/{b}/{b}")
    return "\n".join(lines) + "\n"
