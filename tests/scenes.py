"""Small analytic test scenes built from flat arrays (ngi_scene_create)."""
import numpy as np

import nanogi_amd as ngi


class SceneBuilder:
    def __init__(self):
        self.pos, self.nrm, self.faces, self.tri_prim, self.prims = [], [], [], [], []

    def prim(self, **fields):
        p = ngi.make_primitive(**fields)
        p["tri_begin"] = len(self.faces)
        p["tri_count"] = 0
        self.prims.append(p)
        return len(self.prims) - 1

    def tri(self, pid, a, b, c, n=None):
        a, b, c = (np.asarray(x, dtype=np.float64) for x in (a, b, c))
        if n is None:
            n = np.cross(b - a, c - a)
            n = n / np.linalg.norm(n)
        base = len(self.pos)
        self.pos += [a, b, c]
        self.nrm += [n, n, n]
        self.faces.append([base, base + 1, base + 2])
        self.tri_prim.append(pid)
        self.prims[pid]["tri_count"] += 1

    def rect(self, pid, center, normal, axis, ha, hb):
        """Rectangle whose winding normal equals `normal`."""
        c, n, a = (np.asarray(x, dtype=np.float64) for x in (center, normal, axis))
        n = n / np.linalg.norm(n)
        a = a / np.linalg.norm(a)
        b = np.cross(n, a)
        v = [c - a * ha - b * hb, c + a * ha - b * hb, c + a * ha + b * hb, c - a * ha + b * hb]
        self.tri(pid, v[0], v[1], v[2])
        self.tri(pid, v[0], v[2], v[3])

    def camera(self, eye, center, up=(0, 1, 0), fov=45.0):
        return self.prim(type=ngi.PRIM_E, sensor_type=ngi.SENSOR_PINHOLE, E_We=(1, 1, 1), E_eye=eye,
                         E_center=center, E_up=up, E_fov_deg=fov)

    def scene(self, aspect=1.0):
        return ngi.Scene.create(np.array(self.pos), np.array(self.faces), np.array(self.tri_prim),
                                np.array(self.prims), normals=np.array(self.nrm), aspect=aspect)


def furnace(Le=(1.0, 0.5, 0.25), rho=(0.5, 0.6, 0.7)):
    """Closed unit cube of inward-facing [L, D] walls, camera at the centre.

    Every path has the same contribution sequence, so the radiance of every
    pixel is Le * sum_{k=0}^{m-2} rho^k for -m m (RR off)."""
    B = SceneBuilder()
    walls = [((0.5, 0, 0.5), (0, 1, 0), (1, 0, 0)), ((0.5, 1, 0.5), (0, -1, 0), (1, 0, 0)),
             ((0, 0.5, 0.5), (1, 0, 0), (0, 0, 1)), ((1, 0.5, 0.5), (-1, 0, 0), (0, 0, 1)),
             ((0.5, 0.5, 0), (0, 0, 1), (1, 0, 0)), ((0.5, 0.5, 1), (0, 0, -1), (1, 0, 0))]
    for c, n, a in walls:
        pid = B.prim(type=ngi.PRIM_L | ngi.PRIM_D, light_type=ngi.LIGHT_AREA, Le=Le, D_R=rho)
        B.rect(pid, c, n, a, 0.5, 0.5)
    B.camera((0.5, 0.5, 0.5), (0.5, 0.5, 0.0), fov=60.0)
    return B.scene()


def furnace_radiance(Le, rho, m):
    Le, rho = np.asarray(Le), np.asarray(rho)
    return Le * sum(rho ** k for k in range(m - 1))


def plane_under_light(a=0.5, b=0.3, h=1.0, Le=(4.0, 3.0, 2.0), rho=(0.8, 0.5, 0.3)):
    """Diffuse plane y = 0 under a downward-facing 2a x 2b rectangle at height
    h; a narrow camera looks straight down at the point below its centre."""
    B = SceneBuilder()
    pl = B.prim(type=ngi.PRIM_D, D_R=rho)
    B.rect(pl, (0, 0, 0), (0, 1, 0), (1, 0, 0), 20.0, 20.0)
    li = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=Le)
    B.rect(li, (0, h, 0), (0, -1, 0), (1, 0, 0), a, b)
    B.camera((0, 0.5 * h, 0), (0, 0, 0), up=(0, 0, -1), fov=0.5)
    return B.scene()


def rect_form_factor(a, b, h):
    """Point form factor to a parallel 2a x 2b rectangle centred above it at
    height h: four corner-aligned rectangles (Howell's catalogue, C-11)."""
    x, y = a / h, b / h
    f = (x / np.sqrt(1 + x * x) * np.arctan(y / np.sqrt(1 + x * x)) +
         y / np.sqrt(1 + y * y) * np.arctan(x / np.sqrt(1 + y * y))) / (2 * np.pi)
    return 4 * f


def light_in_view(Le=(3.0, 2.0, 1.0)):
    """A square light facing the camera, covering the centre of the view."""
    B = SceneBuilder()
    li = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=Le)
    B.rect(li, (0, 0, 0), (0, 0, 1), (1, 0, 0), 0.5, 0.5)
    B.camera((0, 0, 3.0), (0, 0, 0), fov=40.0)
    return B.scene()
