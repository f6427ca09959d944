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


def light_in_view(Le=(3.0, 2.0, 1.0), half=0.5):
    """A square light (2 half x 2 half) facing the camera, covering the centre of the view."""
    B = SceneBuilder()
    li = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=Le)
    B.rect(li, (0, 0, 0), (0, 0, 1), (1, 0, 0), half, half)
    B.camera((0, 0, 3.0), (0, 0, 0), fov=40.0)
    return B.scene()


def rect_form_factor_at(x, z, a, b, h):
    """Form factor from a point (x, 0, z) facing up to the parallel 2a x 2b
    rectangle centred above the origin at height h (the point's projection
    inside it): superposition of four corner-aligned rectangles (Howell C-11)."""
    def corner(X, Y):
        X, Y = X / h, Y / h
        return (X / np.sqrt(1 + X * X) * np.arctan(Y / np.sqrt(1 + X * X)) +
                Y / np.sqrt(1 + Y * Y) * np.arctan(X / np.sqrt(1 + Y * Y))) / (2 * np.pi)
    return sum(corner(sx, sz) for sx in (a - x, a + x) for sz in (b - z, b + z))


def plane_under_light_wide(a=0.5, b=0.3, h=1.0, Le=(4.0, 3.0, 2.0), rho=(0.8, 0.5, 0.3), fov=30.0):
    """plane_under_light with a camera at half height looking straight down
    with a wide field of view (its footprint stays under the light)."""
    B = SceneBuilder()
    pl = B.prim(type=ngi.PRIM_D, D_R=rho)
    B.rect(pl, (0, 0, 0), (0, 1, 0), (1, 0, 0), 20.0, 20.0)
    li = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=Le)
    B.rect(li, (0, h, 0), (0, -1, 0), (1, 0, 0), a, b)
    B.camera((0, 0.5 * h, 0), (0, 0, 0), up=(0, 0, -1), fov=fov)
    return B.scene()


def tilted_normals_box(Le=(6.0, 5.0, 4.0)):
    """Open box with a diffuse floor and a glossy back wall whose shading
    normals are tilted away from the geometric ones, one area light (0.4 x
    0.4) under the ceiling, camera in front: exercises the shading-normal
    corrections of both transport directions."""
    B = SceneBuilder()
    fl = B.prim(type=ngi.PRIM_D, D_R=(0.7, 0.6, 0.5))
    n_fl = np.array([0.25, 1.0, 0.1])
    n_fl = n_fl / np.linalg.norm(n_fl)
    v = [np.array(p, dtype=float) for p in ((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1))]
    B.tri(fl, v[0], v[3], v[2], n_fl)
    B.tri(fl, v[0], v[2], v[1], n_fl)
    wall = B.prim(type=ngi.PRIM_G, G_R=(0.9, 0.8, 0.7), G_eta=(0.2, 0.4, 1.4), G_k=(3.9, 2.4, 1.8), G_roughness=0.3)
    n_w = np.array([0.2, -0.15, 1.0])
    n_w = n_w / np.linalg.norm(n_w)
    w = [np.array(p, dtype=float) for p in ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))]
    B.tri(wall, w[0], w[1], w[2], n_w)
    B.tri(wall, w[0], w[2], w[3], n_w)
    side = B.prim(type=ngi.PRIM_D, D_R=(0.3, 0.5, 0.4))
    B.rect(side, (0, 0.5, 0.5), (1, 0, 0), (0, 0, 1), 0.5, 0.5)
    li = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=Le)
    B.rect(li, (0.5, 0.95, 0.5), (0, -1, 0), (1, 0, 0), 0.2, 0.2)
    B.camera((0.5, 0.5, 2.0), (0.5, 0.4, 0.0), fov=50.0)
    return B.scene()


def plane_with_light(kind, h=1.0, Le=(2.0, 1.5, 1.0), rho=(0.8, 0.5, 0.3), fov=20.0, half=20.0):
    """Diffuse plane y = 0 (2 half x 2 half) lit by a point light at (0, h, 0)
    or a directional light shining straight down; camera at h/2 looking down."""
    B = SceneBuilder()
    pl = B.prim(type=ngi.PRIM_D, D_R=rho)
    B.rect(pl, (0, 0, 0), (0, 1, 0), (1, 0, 0), half, half)
    if kind == "point":
        B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_POINT, Le=Le, light_position=(0, h, 0))
    else:
        B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_DIRECTIONAL, Le=Le, light_direction=(0, -1, 0))
    B.camera((0, 0.5 * h, 0), (0, 0, 0), up=(0, 0, -1), fov=fov)
    return B.scene()


def pixel_plane_points(W, h, fov, sub=1):
    """(x, z) of the plane points seen through the downward camera of
    plane_under_light_wide / plane_with_light: `sub` x `sub` evenly spaced
    raster points per pixel (raster is linear in x and z on the plane, so
    their mean approximates the pixel average to O(1/sub^2))."""
    t = np.tan(np.radians(fov * 0.5))
    c = (np.arange(W * sub) + 0.5) / (W * sub) * 2 - 1
    X, Z = np.meshgrid(c * t * 0.5 * h, c * t * 0.5 * h)
    return X.ravel(), Z.ravel()


def _cube(B, pid, lo, hi, inward):
    """Six faces of the axis-aligned box [lo, hi] with outward (or inward) winding normals."""
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    c, h = (lo + hi) * 0.5, (hi - lo) * 0.5
    sgn = -1.0 if inward else 1.0
    for ax in range(3):
        a1, a2 = (ax + 1) % 3, (ax + 2) % 3
        for side in (-1.0, 1.0):
            n = np.zeros(3)
            n[ax] = side * sgn
            ctr = c.copy()
            ctr[ax] += side * h[ax]
            axis = np.zeros(3)
            axis[a1] = 1.0
            # rect() spans `axis` by ha and cross(n, axis) by hb
            hb = h[a2]
            B.rect(pid, ctr, n, axis, h[a1], hb)


def glass_cube_in_light_box(spec_type, eta=1.5, Le=(2.0, 1.0, 0.5), look=(0.0, 0.0, -1.0), fov=30.0):
    """Camera at the centre of a unit glass cube (S primitive, Eta1 = 1
    outside / Eta2 = eta inside, outward normals) inside a closed box of
    3 x 3 x 3 inward-facing [L]-only walls.  With -m 3 a camera ray that
    leaves the glass reaches a wall: pt sees eta^2 Le times the probability
    of transmission (radiance scales as n^2 across an interface)."""
    B = SceneBuilder()
    g = B.prim(type=ngi.PRIM_S, spec_type=spec_type, S_R=(1, 1, 1), S_eta1=1.0, S_eta2=eta)
    _cube(B, g, (0, 0, 0), (1, 1, 1), inward=False)
    li = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=Le)
    _cube(B, li, (-1, -1, -1), (2, 2, 2), inward=True)
    look = np.asarray(look, dtype=np.float64)
    B.camera((0.5, 0.5, 0.5), tuple(np.array([0.5, 0.5, 0.5]) + look), fov=fov)
    return B.scene()


def slab_furnace(eta=1.5, Le=(1.0, 0.7, 0.4)):
    """Closed [L]-only box (2 x 2 x 2, inward normals) with a Fresnel glass slab
    (1.2 x 0.3 x 1.2) in the middle and the camera above it looking down.
    Entering and leaving the slab scale radiance by (1/eta)^2 and eta^2, and a
    reflection's Fr / pdf is 1, so every pt path that reaches a wall carries
    exactly Le (paths cut off by -m carry 0)."""
    B = SceneBuilder()
    li = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=Le)
    _cube(B, li, (-1, -1, -1), (1, 1, 1), inward=True)
    g = B.prim(type=ngi.PRIM_S, spec_type=ngi.SPEC_FRESNEL, S_R=(1, 1, 1), S_eta1=1.0, S_eta2=eta)
    _cube(B, g, (-0.6, -0.15, -0.6), (0.6, 0.15, 0.6), inward=False)
    B.camera((0.1, 0.8, 0.2), (0.0, 0.0, 0.0), up=(0, 0, -1), fov=60.0)
    return B.scene()


def glass_room(spec_type=None, eta=1.5, Le=(6.0, 5.0, 4.0), rho=(0.7, 0.5, 0.3)):
    """A unit glass box (outward normals) holding the camera and a diffuse
    floor; one 0.4 x 0.4 area light outside, above the box.  Light reaches
    the floor only through the glass top, so path tracing (EL: eta^2 on
    leaving the glass) and light tracing (LE: no scaling) meet only if both
    transport conventions are right."""
    B = SceneBuilder()
    g = B.prim(type=ngi.PRIM_S, spec_type=spec_type if spec_type is not None else ngi.SPEC_REFRACTION,
               S_R=(1, 1, 1), S_eta1=1.0, S_eta2=eta)
    _cube(B, g, (0, 0, 0), (1, 1, 1), inward=False)
    fl = B.prim(type=ngi.PRIM_D, D_R=rho)
    B.rect(fl, (0.5, 0.05, 0.5), (0, 1, 0), (1, 0, 0), 0.49, 0.49)
    li = B.prim(type=ngi.PRIM_L, light_type=ngi.LIGHT_AREA, Le=Le)
    B.rect(li, (0.5, 1.6, 0.5), (0, -1, 0), (1, 0, 0), 0.2, 0.2)
    B.camera((0.5, 0.9, 0.9), (0.5, 0.05, 0.4), fov=50.0)
    return B.scene()
