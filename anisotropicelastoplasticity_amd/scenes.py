"""Seeded scene builders for the benchmark configurations and the parity cases.

Each builder returns a `Scene`: simulation parameters, grid, level-set primitives and the initial
particle arrays in the planar layout of the C ABI. Positions come from the seeded stratified-jitter
generator of the C ABI (mpm_generate_block), so identical seeds give identical inputs to the device
and to any CPU restatement. These replace the unseeded factories of the reference
(ParticleSystem.cpp:119-401); material defaults are those of the configurations named below.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .api import generate_block, generate_shape, levelset


@dataclass
class Scene:
    name: str
    params: object
    gmin: tuple
    gmax: tuple
    res: tuple
    prims: list
    particles: dict
    dt: float
    substeps: int
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.particles["x"].shape[1]

    def load_into(self, solver):
        solver.set_grid(self.gmin, self.gmax, self.res)
        solver.set_params(self.params)
        solver.set_levelset(self.prims)
        p = self.particles
        solver.set_particles(p["x"], p["mass"], v=p["v"], B=p.get("B"), F=p.get("F"), plastic=p["plastic"],
                             volume=p["volume"], phi=p.get("phi"))


def _particles(x, rho, vol_total, v0, material, phi=None):
    n = x.shape[1]
    vp = vol_total / n
    v = np.empty((3, n), np.float32)
    v[:] = np.asarray(v0, np.float32).reshape(3, 1)
    F = np.zeros((9, n), np.float32)
    F[0] = F[4] = F[8] = 1.0
    return dict(x=x, v=v, B=np.zeros((9, n), np.float32), F=F,
                plastic=np.full(n, 1.0 if material == _lib.SNOW else 0.0, np.float32),
                mass=np.full(n, rho * vp, np.float32), volume=np.full(n, vp, np.float32),
                phi=phi)


def _shape_particles(x, total_mass, total_volume, v0, material):
    """The reference factories' per-particle mass total_mass / n (ParticleSystem.cpp:149-153);
    their volumes are placeholders (ones) that the first P2G's density estimate replaces
    (HybridSolver.cpp:242-249), so here V = total_volume / n until mpm_estimate_volumes."""
    n = x.shape[1]
    p = _particles(x, 0.0, total_volume, v0, material)
    p["mass"] = np.full(n, total_mass / n, np.float32)
    return p


def snow_ball(seed, center, radius, n):
    """ParticleSystem::SnowBall (ParticleSystem.cpp:119-178): n particles uniform in the ball, at
    rest, total mass 100 * 3.14 * r^3 (the reference's constant), snow E = 1.4e5, nu = 0.2."""
    x = generate_shape(seed, _lib.SHAPE_BALL, list(center) + [radius], n)
    return _shape_particles(x, 100.0 * 3.14 * radius ** 3, 4.0 / 3.0 * np.pi * radius ** 3, (0.0, 0.0, 0.0), _lib.SNOW)


def sand_ball(seed, center, radius, n):
    """ParticleSystem::SandBall (ParticleSystem.cpp:180-238): as snow_ball with total mass
    1300 * 3.14 * r^3; sand E = 3.537e5, nu = 0.3."""
    x = generate_shape(seed, _lib.SHAPE_BALL, list(center) + [radius], n)
    return _shape_particles(x, 1300.0 * 3.14 * radius ** 3, 4.0 / 3.0 * np.pi * radius ** 3, (0.0, 0.0, 0.0), _lib.SAND)


def sand_block(seed, bmin, bmax, hole_radius, n):
    """ParticleSystem::SandBlock (ParticleSystem.cpp:240-323): n particles uniform in the box minus
    the ball of radius hole_radius around (min x, min y, mid z); total mass
    1300 * (box volume - 0.25 pi r^3), as the reference computes it."""
    x = generate_shape(seed, _lib.SHAPE_BLOCK, list(bmin) + list(bmax) + [hole_radius], n)
    vol = float(np.prod(np.subtract(bmax, bmin))) - 0.25 * np.pi * hole_radius ** 3
    return _shape_particles(x, 1300.0 * vol, vol, (0.0, 0.0, 0.0), _lib.SAND)


def sand_cylinder(seed, base, radius, height, n, axis=2):
    """ParticleSystem::SandCylinder (ParticleSystem.cpp:325-401): n particles uniform in the
    cylinder over the disk of radius r at `base`, height along `axis` (z in the reference), initial
    velocity -1 along the axis, total mass 1300 pi r^2 height."""
    x = generate_shape(seed, _lib.SHAPE_CYLINDER, list(base) + [radius, height, axis], n)
    v0 = [0.0, 0.0, 0.0]
    v0[axis] = -1.0
    vol = np.pi * radius ** 2 * height
    return _shape_particles(x, 1300.0 * vol, vol, v0, _lib.SAND)


def snow_params(order=_lib.KERNEL_QUADRATIC, gravity=(0.0, -9.8, 0.0)):
    p = _lib.default_params()
    p.kernel_order = order
    p.material = _lib.SNOW
    p.stress_at_predicted_F = 0
    p.friction_model = _lib.FRICTION_STICKY
    for i in range(3):
        p.gravity[i] = gravity[i]
    p.youngs_modulus = 1.4e5
    p.poisson_ratio = 0.2
    p.critical_compression = 2.5e-2
    p.critical_stretch = 7.5e-3
    p.hardening = 10.0
    return p


def sand_params(order=_lib.KERNEL_QUADRATIC, E=3.5e5, nu=0.3, phi_deg=30.0, mu=0.4, gravity=(0.0, -9.8, 0.0)):
    p = _lib.default_params()
    p.kernel_order = order
    p.material = _lib.SAND
    p.stress_at_predicted_F = 0
    p.friction_model = _lib.FRICTION_COULOMB
    p.friction = mu
    for i in range(3):
        p.gravity[i] = gravity[i]
    p.youngs_modulus = E
    p.poisson_ratio = nu
    p.dp_h[0] = phi_deg  # constant friction angle: h1 = h3 = 0
    p.dp_h[1] = 0.0
    p.dp_h[2] = 0.0
    p.dp_h[3] = 0.0
    return p


def config1_snow_block(seed=1, y_offset=0.0, substeps=200):
    """Plain snow block: grid 64^3 over the unit cube, 32,768 particles (8 per cell) in
    [0.3,0.55]^3, v0 = (0,-1,0), fixed-corotated snow E=1.4e5 nu=0.2 rho=400, theta_c=2.5e-2,
    theta_s=7.5e-3, xi=10, gravity -9.8 y, dt=1e-4, sticky ground plane y=0.05.
    `y_offset` lowers the block (a variant that reaches the ground within the run)."""
    lo = (0.3, 0.3 - y_offset, 0.3)
    hi = (0.55, 0.55 - y_offset, 0.55)
    x = generate_block(seed, lo, hi, (16, 16, 16), 2)
    parts = _particles(x, 400.0, 0.25 ** 3, (0.0, -1.0, 0.0), _lib.SNOW)
    prims = [levelset(_lib.LS_PLANE, (0.0, 1.0, 0.0, 0.05))]
    return Scene("config1_snow_block", snow_params(), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (64, 64, 64), prims, parts,
                 1e-4, substeps, dict(ppc=8))


def config2_sand_column(seed=2, substeps=1000, cells=(32, 128, 32)):
    """Drucker-Prager sand column collapse: grid 128^3, 2^20 particles in [0.4,0.6]x[0.05,0.65]x
    [0.4,0.6] (a 32x128x32 jittered lattice with 8 particles per lattice cell, about 20.8 per grid
    cell), E=3.5e5 nu=0.3 rho=2200, constant friction angle 30 deg, cohesion 0; box walls
    [0.05,0.95]^3 plus a static sphere (r=0.1, centre (0.5,0.15,0.7)), Coulomb friction mu=0.4,
    dt=5e-5."""
    lo, hi = (0.4, 0.05, 0.4), (0.6, 0.65, 0.6)
    x = generate_block(seed, lo, hi, cells, 2)
    vol = 0.2 * 0.6 * 0.2
    parts = _particles(x, 2200.0, vol, (0.0, 0.0, 0.0), _lib.SAND)
    prims = [levelset(_lib.LS_BOX, (0.05, 0.05, 0.05, 0.0), (0.95, 0.95, 0.95, 0.0)),
             levelset(_lib.LS_SPHERE, (0.5, 0.15, 0.7, 0.1))]
    n = x.shape[1]
    return Scene("config2_sand_column", sand_params(), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (128, 128, 128), prims, parts,
                 5e-5, substeps, dict(ppc=n / (vol * 128 ** 3)))


def particle_friction_angles(seed, n, lo=25.0, hi=40.0):
    """Per-particle friction angles (degrees) uniform in [lo, hi): for particle id,
    u = (splitmix64((seed * 0x9E3779B97F4A7C15) ^ id) >> 40) * 2^-24 (the counter hash of the position
    generator with another key), angle = lo + (hi - lo) u."""
    m64 = (1 << 64) - 1
    z = (np.arange(n, dtype=np.uint64) ^ np.uint64((seed * 0x9E3779B97F4A7C15) & m64))
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)
    return (lo + (hi - lo) * u).astype(np.float32)


def config4_sand_phi(seed=4, substeps=1000, cells=(64, 128, 64)):
    """The surrogate's workload: 2^22 Drucker-Prager sand particles (config 2's material) with a
    per-particle friction angle uniform in [25, 40) degrees, on a 256^3 grid: a 64x128x64 jittered
    lattice of grid-cell-sized lattice cells (8 particles per grid cell) filling the column
    [0.375,0.625]x[0.05,0.55]x[0.375,0.625]; box walls
    [0.05,0.95]^3 plus a static sphere (r=0.1, centre (0.5,0.15,0.7)), Coulomb friction 0.4,
    dt=5e-5 (CFL about 0.19 at dx=1/256)."""
    lo, hi = (0.375, 0.05, 0.375), (0.625, 0.55, 0.625)
    x = generate_block(seed, lo, hi, cells, 2)
    vol = float(np.prod(np.subtract(hi, lo)))
    n = x.shape[1]
    parts = _particles(x, 2200.0, vol, (0.0, 0.0, 0.0), _lib.SAND, phi=particle_friction_angles(seed, n))
    p = sand_params()
    p.per_particle_phi = 1
    prims = [levelset(_lib.LS_BOX, (0.05, 0.05, 0.05, 0.0), (0.95, 0.95, 0.95, 0.0)),
             levelset(_lib.LS_SPHERE, (0.5, 0.15, 0.7, 0.1))]
    return Scene("config4_sand_phi", p, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (256, 256, 256), prims, parts, 5e-5,
                 substeps, dict(ppc=n / (vol * 256 ** 3)))


def surrogate_training_column(seed=41, substeps=1000, cells=(32, 128, 32)):
    """The run whose exact return map supplies the surrogate's training pairs: config 2's sand column
    (2^20 particles, 128^3, dt 5e-5, same walls and sphere) with config 4's per-particle friction
    angles uniform in [25, 40) degrees, from its own seed."""
    sc = config2_sand_column(seed=seed, substeps=substeps, cells=cells)
    sc.params.per_particle_phi = 1
    sc.particles["phi"] = particle_friction_angles(seed, sc.n)
    sc.name = "surrogate_training_column"
    return sc


def small_scene(material, order, predicted, friction_model, seed=7, res=(24, 24, 24), cells=(6, 5, 6), ppa=2,
                lo=(0.3, 0.2, 0.35), hi=(0.55, 0.42, 0.6), v0=(0.3, -1.5, 0.2), z_up=False, prims=None,
                dt=2e-4, E=None):
    """A small scene for per-stage parity: any material / kernel / stress mode / friction model."""
    grav = (0.0, 0.0, -9.8) if z_up else (0.0, -9.8, 0.0)
    if material == _lib.SNOW:
        p = snow_params(order, grav)
        rho = 400.0
    else:
        p = sand_params(order, gravity=grav)
        rho = 2200.0
    if E is not None:
        p.youngs_modulus = E
    p.stress_at_predicted_F = int(predicted)
    p.friction_model = friction_model
    if friction_model == _lib.FRICTION_REFERENCE:
        p.friction = 0.2
    x = generate_block(seed, lo, hi, cells, ppa)
    vol = float(np.prod(np.subtract(hi, lo)))
    parts = _particles(x, rho, vol, v0, material)
    if prims is None:
        if z_up:
            prims = [levelset(_lib.LS_PLANE, (0.0, 0.0, 1.0, 0.2))]
        else:
            prims = [levelset(_lib.LS_PLANE, (0.0, 1.0, 0.0, 0.18))]
    return Scene("small", p, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), res, prims, parts, dt, 1)


def batch_scene(k, rng, n_range=(20000, 500000), res_choices=(64, 80, 96, 112, 128), dt=5e-5, substeps=500):
    """Scene k of a batched run: snow or sand (50/50), a grid of 64^3..128^3 over the unit cube, a
    particle count log-uniform in n_range, E log-uniform in [1e5, 1e6], nu uniform in [0.2, 0.35],
    sand friction angle uniform in [25, 40] degrees. The particles form a cube of c^3 lattice cells
    (8 jittered particles per cell, one lattice cell per grid cell) resting 1 cm above the floor of the
    box walls [0.05, 0.95]^3 (Coulomb friction 0.4) and moving into it at 1 m/s, so that the blocks
    hit the floor within the first substeps and sand flows plastically."""
    sand = bool(rng.random() < 0.5)
    r = int(rng.choice(res_choices))
    n_want = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
    E = float(np.exp(rng.uniform(np.log(1e5), np.log(1e6))))
    nu = float(rng.uniform(0.2, 0.35))
    phi = float(rng.uniform(25.0, 40.0))
    c = max(2, int(round((n_want / 8.0) ** (1.0 / 3.0))))
    edge = c / r
    lo = (0.5 - edge / 2, 0.06, 0.5 - edge / 2)
    hi = (lo[0] + edge, lo[1] + edge, lo[2] + edge)
    x = generate_block(1000 + k, lo, hi, (c, c, c), 2)
    if sand:
        params = sand_params(E=E, nu=nu, phi_deg=phi)
        rho = 2200.0
    else:
        params = snow_params()
        params.youngs_modulus = E
        params.poisson_ratio = nu
        params.friction_model = _lib.FRICTION_COULOMB
        rho = 400.0
    params.friction = 0.4
    parts = _particles(x, rho, edge ** 3, (0.0, -1.0, 0.0), _lib.SAND if sand else _lib.SNOW)
    prims = [levelset(_lib.LS_BOX, (0.05, 0.05, 0.05, 0.0), (0.95, 0.95, 0.95, 0.0))]
    return Scene(f"batch{k}", params, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (r, r, r), prims, parts, dt, substeps,
                 dict(sand=sand, E=E, nu=nu, phi=phi if sand else None))


def config5_batch(seed=5, nscenes=64, **kw):
    """Batched many-scene run: `nscenes` independent scenes drawn by batch_scene from one seed
    (64 scenes: about 9.5 million particles), dt = 5e-5, 500 substeps."""
    rng = np.random.default_rng(seed)
    return [batch_scene(k, rng, **kw) for k in range(nscenes)]


def grid_mesh(nx, nz, edge, origin, y):
    """A flat square triangle mesh in the plane y = const: nx x nz vertices with spacing `edge`
    starting at `origin` (x, z), two triangles per quad (counter-clockwise seen from +y, so the
    normals point up). Returns x (3, nv) and faces (nf, 3)."""
    i, k = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    x = np.stack([origin[0] + edge * i.ravel(), np.full(nx * nz, float(y)), origin[1] + edge * k.ravel()]).astype(np.float32)
    vid = (i * nz + k)
    a, b = vid[:-1, :-1].ravel(), vid[1:, :-1].ravel()
    c, d = vid[1:, 1:].ravel(), vid[:-1, 1:].ravel()
    faces = np.concatenate([np.stack([a, d, b], 1), np.stack([b, d, c], 1)]).astype(np.int32)
    return x, faces


def config3_cloth_sand(seed=3, n_side=256, edge=1.0 / 300.0, sand_cells=(40, 39, 40), substeps=2000):
    """Cloth-sand coupling: a 256 x 256-vertex cloth (130,050 triangles, edge 1/300) flat at y = 0.7
    above a static sphere (r = 0.15, centre (0.5, 0.5, 0.5)), E = 200, nu = 0.3, rho = 2e3, thickness
    0.04 (the reference's main.cpp:78 defaults), normal stiffness k = 1e4, shear stiffness gamma = 1e3,
    cloth friction angle 20 deg; about 500k Drucker-Prager sand particles (config 2's material) in a
    block [0.35, 0.65] x [0.72, 0.92] x [0.35, 0.65] falling onto it; grid 128^3, box walls
    [0.05, 0.95]^3, Coulomb friction 0.4, dt = 2e-5. Returns (scene, mesh dict)."""
    lo, hi = (0.35, 0.72, 0.35), (0.65, 0.92, 0.65)
    x = generate_block(seed, lo, hi, sand_cells, 2)
    vol = float(np.prod(np.subtract(hi, lo)))
    parts = _particles(x, 2200.0, vol, (0.0, 0.0, 0.0), _lib.SAND)
    prims = [levelset(_lib.LS_BOX, (0.05, 0.05, 0.05, 0.0), (0.95, 0.95, 0.95, 0.0)),
             levelset(_lib.LS_SPHERE, (0.5, 0.5, 0.5, 0.15))]
    side = edge * (n_side - 1)
    mx, faces = grid_mesh(n_side, n_side, edge, (0.5 - side / 2, 0.5 - side / 2), 0.7)
    mesh = dict(x=mx, faces=faces, youngs_modulus=200.0, poisson_ratio=0.3, shear_stiffness=1e3, stiffness=1e4,
                friction_angle=20.0, density=2e3, thickness=0.04)
    sc = Scene("config3_cloth_sand", sand_params(), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (128, 128, 128), prims, parts,
               2e-5, substeps, dict(mesh=mesh))
    return sc, mesh
