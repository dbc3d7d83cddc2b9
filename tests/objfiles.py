"""Small Wavefront OBJ / MTL scenes written into a directory for the ingest
tests (the reference ships no OBJ assets)."""
from __future__ import annotations

from pathlib import Path

from tracer_amd import abi


def write(dirpath: Path, name: str, obj: str, mtl: str | None = None, mtl_name: str = "scene.mtl") -> Path:
    dirpath.mkdir(parents=True, exist_ok=True)
    if mtl is not None:
        (dirpath / mtl_name).write_text(mtl)
    p = dirpath / name
    p.write_text(obj)
    return p


def camera(pos, target, up=(0.0, 1.0, 0.0), fov_deg=40.0, aspect=1.0) -> abi.TrCamera:
    import math
    c = abi.TrCamera()
    c.pos = abi.f3(*pos)
    c.target = abi.f3(*target)
    c.up = abi.f3(*up)
    c.fov_y = math.radians(fov_deg)
    c.lens_radius = 0.0
    c.focal_distance = 1.0
    c.aspect = aspect
    return c


def box_room_obj(light_ke=(12.0, 12.0, 12.0), block_ks=(0.35, 0.35, 0.35), block_ns=25.0,
                 glass_block: bool = False, texture: str | None = None, closed: bool = False) -> tuple[str, str]:
    """A closed-on-five-sides room [-1, 1]^3 open towards +z (closed: all six), a ceiling quad
    light facing down, a tall block with a Phong specular lobe (or of glass,
    Ni 1.5) and a short diffuse block; faces are quads (split by the reader),
    normals given per face with v//vn corners.  With `texture` (an image file
    name next to the MTL) the floor is a diffuse and the back wall a glossy
    material whose map_Kd is that image, with v/vt/vn corners whose texture
    coordinates run past 1 (the lookup wraps)."""
    mtl = f"""# materials
newmtl white
Kd 0.73 0.73 0.73
newmtl red
Kd 0.65 0.05 0.05
newmtl green
Kd 0.12 0.45 0.15
newmtl glossy
Kd 0.45 0.40 0.30
Ks {block_ks[0]} {block_ks[1]} {block_ks[2]}
Ns {block_ns}
newmtl glass
Ks 0.95 0.95 0.95
Ni 1.5
newmtl light
Kd 0.78 0.78 0.78
Ke {light_ke[0]} {light_ke[1]} {light_ke[2]}
"""
    if texture:
        mtl += f"""newmtl tex_floor
Kd 0.9 0.9 0.9
map_Kd {texture}
newmtl tex_wall
Kd 0.8 0.7 0.6
Ks 0.3 0.3 0.3
Ns 15
map_Kd {texture}
"""
    lines = ["mtllib scene.mtl"]
    v, n, vt = [], [], []

    def quad(c0, c1, c2, c3, normal, mat, group=None, uv=None):
        if group:
            lines.append(f"o {group}")
        lines.append(f"usemtl {mat}")
        base = len(v)
        v.extend([c0, c1, c2, c3])
        n.append(normal)
        ni = len(n)
        if uv is None:
            lines.append("f " + " ".join(f"{base + k + 1}//{ni}" for k in range(4)))
        else:
            tb = len(vt)
            vt.extend(uv)
            lines.append("f " + " ".join(f"{base + k + 1}/{tb + k + 1}/{ni}" for k in range(4)))

    # room: floor, ceiling, back wall, left (red), right (green); normals point inwards
    floor_uv = ((0.0, 0.0), (2.5, 0.0), (2.5, 2.5), (0.0, 2.5)) if texture else None
    wall_uv = ((-0.25, 0.0), (1.5, 0.0), (1.5, 1.0), (-0.25, 1.0)) if texture else None
    quad((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1), (0, 1, 0), "tex_floor" if texture else "white", "room",
         floor_uv)
    quad((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1), (0, -1, 0), "white")
    quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (0, 0, 1), "tex_wall" if texture else "white", None,
         wall_uv)
    quad((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (1, 0, 0), "red")
    quad((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1), (-1, 0, 0), "green")
    if closed:   # the sixth wall, at z = +1 facing -z
        quad((1, -1, 1), (-1, -1, 1), (-1, 1, 1), (1, 1, 1), (0, 0, -1), "white")
    # light: a quad just below the ceiling facing down
    y = 0.998
    quad((-0.25, y, -0.25), (0.25, y, -0.25), (0.25, y, 0.25), (-0.25, y, 0.25), (0, -1, 0), "light", "light")

    def block(x0, x1, z0, z1, h, mat, name):
        y0 = -0.998
        first = True
        for c, nn in (
            (((x0, h, z1), (x1, h, z1), (x1, h, z0), (x0, h, z0)), (0, 1, 0)),
            (((x0, y0, z1), (x1, y0, z1), (x1, h, z1), (x0, h, z1)), (0, 0, 1)),
            (((x1, y0, z0), (x0, y0, z0), (x0, h, z0), (x1, h, z0)), (0, 0, -1)),
            (((x0, y0, z0), (x0, y0, z1), (x0, h, z1), (x0, h, z0)), (-1, 0, 0)),
            (((x1, y0, z1), (x1, y0, z0), (x1, h, z0), (x1, h, z1)), (1, 0, 0)),
        ):
            quad(*c, nn, mat, name if first else None)
            first = False

    block(-0.65, -0.05, -0.7, -0.1, 0.25, "glass" if glass_block else "glossy", "tall")
    block(0.1, 0.6, 0.05, 0.55, -0.4, "white", "short")
    text = "\n".join([f"v {a} {b} {c}" for a, b, c in v] + [f"vn {a} {b} {c}" for a, b, c in n] +
                     [f"vt {a} {b}" for a, b in vt] + lines) + "\n"
    return text, mtl
