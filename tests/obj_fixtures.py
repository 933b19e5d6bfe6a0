"""Small OBJ/MTL files written by the tests (a cube with two materials, uvs
and normals; a pentagon; files with faults)."""
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
