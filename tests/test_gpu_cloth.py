"""GPU parity of the cloth (LagrangianMesh) path through the C ABI against the FP64 oracle.

Tolerances (FP32 device state against the FP64 restatement):
  one substep, per stage    grid m, momentum (cloth forces included), v, v_BF, mesh vertices and
                            directions D, sand: each error at most 2x the FP32 oracle's own error
                            against the FP64 oracle, plus 1e-6 (the in-plane forces cancel to small
                            strains, which FP32 positions resolve to a few percent)
  60 substeps, sphere       mesh vertices |dx| <= 1e-4 h, D relative L2 <= 1e-3, with the return map
                            active (some elements changed by it)
  fixed vertices            fixed vertices keep their positions exactly; parity as above
  cloth alone               no particles in the context: parity as above; a second mpm_set_mesh
                            replaces the cloth and repeats the run bit for bit
  full size (256^2 cloth, about 500k sand particles, 128^3 grid), 200 substeps: before any contact
                            the cloth falls as a rigid body (v_y = n dt g within 1e-4 relative, no
                            in-plane velocity) and its directions keep their rest values within 1e-4
"""
import sys

import numpy as np
import pytest

from anisotropicelastoplasticity_amd import Solver, _lib, api, levelset, scenes
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def cloth_scene(n_side=20, fixed_corners=False, stretch=True, seed=5):
    """A 20 x 20 cloth (edge 1/48) just above a sphere, a small sand block above the cloth, and a
    velocity field that stretches, shears and bends the cloth so every force term is non-zero."""
    prims = [levelset(_lib.LS_BOX, (0.05, 0.05, 0.05, 0.0), (0.95, 0.95, 0.95, 0.0)),
             levelset(_lib.LS_SPHERE, (0.5, 0.42, 0.5, 0.15))]
    sc = scenes.small_scene(_lib.SAND, 2, 0, _lib.FRICTION_COULOMB, seed=seed, res=(32, 32, 32), cells=(3, 2, 3),
                            lo=(0.44, 0.66, 0.44), hi=(0.56, 0.72, 0.56), v0=(0.0, -1.0, 0.0), prims=prims, dt=1e-4)
    edge = 1.0 / 48.0
    side = edge * (n_side - 1)
    x, faces = scenes.grid_mesh(n_side, n_side, edge, (0.5 - side / 2, 0.5 - side / 2), 0.6)
    v = np.zeros_like(x)
    v[1] = -0.8
    if stretch:
        v[0] = (x[0] - 0.5) * 2.0
        v[2] = -(x[2] - 0.5) * 1.5
        v[1] += 0.3 * np.sin(25.0 * x[0]) * np.cos(18.0 * x[2])
    fixed = None
    if fixed_corners:
        fixed = np.zeros(x.shape[1], np.uint8)
        fixed[[0, n_side - 1, n_side * (n_side - 1), n_side * n_side - 1]] = 1
    return sc, x, v.astype(np.float32), faces, fixed


def _both(sc, x, v, faces, fixed):
    o = Oracle(sc.params, sc.gmin, sc.gmax, sc.res, prims=sc.prims)
    o.load_scene(sc)
    o.set_mesh(x, faces, v=v, fixed=fixed)
    s = Solver()
    sc.load_into(s)
    s.set_mesh(x, faces, v=v, fixed=fixed)
    return s, o


def _oracle_substep_then_stages(sc, x, v, faces, fixed, precision):
    o = Oracle(sc.params, sc.gmin, sc.gmax, sc.res, prims=sc.prims, precision=precision)
    o.load_scene(sc)
    o.set_mesh(x, faces, v=v, fixed=fixed)
    o.substep(sc.dt)
    o.stage(_lib.STAGE_P2G, sc.dt)
    o.stage(_lib.STAGE_GRID, sc.dt)
    g = o.get_grid()
    o.stage(_lib.STAGE_G2P, sc.dt)
    return g, o.get_mesh(), o.get_particles()


def _errors(g, m, q, g64, m64, q64, dt, h):
    on = g64["m"] > 0
    mom = g64["p"] + dt * g64["f"]
    gm = g["mom"] if "mom" in g else g["p"] + dt * g["f"]
    return dict(m=rel(g["m"], g64["m"]), mom=rel(gm[:, on], mom[:, on]), v=rel(g["v"][:, on], g64["v"][:, on]),
                vbf=rel(g["vbf"][:, on], g64["vbf"][:, on]), mesh_x=np.max(np.abs(m["vx"] - m64["vx"])) / h,
                mesh_v=rel(m["vv"], m64["vv"]), D=rel(m["D"], m64["D"]),
                sand_x=np.max(np.abs(q["x"] - q64["x"])) / h, sand_v=rel(q["v"], q64["v"]))


def test_cloth_one_substep_stages():
    """The in-plane forces come from F - R with strains of order 1e-4, formed from FP32 vertex
    positions: FP32 rounding alone moves them by a few percent and the grid velocity by about 1e-4
    relative. The bar is therefore the FP32 oracle's own distance from the FP64 oracle (same
    algorithm, FP32 arithmetic): the GPU may be at most twice as far, plus 1e-6."""
    sc, x, v, faces, fixed = cloth_scene()
    dt = sc.dt
    h = 1.0 / sc.res[0]
    say(f"[cloth stages] particles {sc.n}, vertices {x.shape[1]}, faces {len(faces)}")
    g64, m64, q64 = _oracle_substep_then_stages(sc, x, v, faces, fixed, 64)
    g32, m32, q32 = _oracle_substep_then_stages(sc, x, v, faces, fixed, 32)
    s = Solver()
    sc.load_into(s)
    s.set_mesh(x, faces, v=v, fixed=fixed)
    with s:
        # one full substep first, so the directions D carry deformation into the next forces
        s.substep(dt)
        for st in range(4):
            s.stage(st, dt)
            if st == _lib.STAGE_GRID:
                g = s.get_grid()
        m, q = s.get_mesh(), s.get_particles()
        stats = s.stats()
    on = g64["m"] > 0
    assert np.array_equal(g["m"] > 0, on), "active node sets differ"
    assert np.max(np.abs(g64["f"][:, on])) > 1.0  # cloth forces are present
    e_gpu = _errors(g, m, q, g64, m64, q64, dt, h)
    e_f32 = _errors(g32, m32, q32, g64, m64, q64, dt, h)
    for k in e_gpu:
        say(f"  {k:7s} gpu {e_gpu[k]:.3e}  fp32 oracle {e_f32[k]:.3e}")
    for k in e_gpu:
        assert e_gpu[k] <= 2.0 * e_f32[k] + 1e-6, k
    assert stats["status"] == 0, stats


@pytest.mark.parametrize("fixed_corners", [False, True])
def test_cloth_drape_60_substeps(fixed_corners):
    sc, x, v, faces, fixed = cloth_scene(fixed_corners=fixed_corners, stretch=False)
    n = 60
    say(f"[cloth drape] fixed={fixed_corners} substeps={n}")
    s, o = _both(sc, x, v, faces, fixed)
    with s:
        s.run(sc.dt, n)
        m = s.get_mesh()
        q = s.get_particles()
        st = s.stats()
    o.run(sc.dt, n)
    mr, qr = o.get_mesh(), o.get_particles()
    h = 1.0 / sc.res[0]
    dxm = np.max(np.abs(m["vx"] - mr["vx"]))
    say(f"  mesh |dx|/h {dxm / h:.3e}, D rel {rel(m['D'], mr['D']):.3e}, sand |dx|/h "
        f"{np.max(np.abs(q['x'] - qr['x'])) / h:.3e}")
    # the sphere stopped part of the cloth (free fall ends at v_y = -0.86) and bent its normals
    assert np.max(mr["vv"][1]) > -0.7
    assert np.max(np.abs(mr["D"][5] - 1.0)) > 1e-2
    assert dxm <= 1e-4 * h
    assert rel(m["D"], mr["D"]) < 1e-3
    assert np.max(np.abs(q["x"] - qr["x"])) <= 1e-4 * h
    if fixed_corners:
        idx = np.flatnonzero(fixed)
        assert np.array_equal(m["vx"][:, idx], x[:, idx])
    assert st["status"] == 0, st


def test_config3_full_size_free_fall():
    sc, mesh = scenes.config3_cloth_sand()
    n = 200
    say(f"[config3] sand {sc.n}, vertices {mesh['x'].shape[1]}, faces {len(mesh['faces'])}, substeps {n}")
    with Solver() as s:
        sc.load_into(s)
        kw = {k: mesh[k] for k in ("youngs_modulus", "poisson_ratio", "shear_stiffness", "stiffness",
                                   "friction_angle", "density", "thickness")}
        s.set_mesh(mesh["x"], mesh["faces"], **kw)
        m0 = s.get_mesh()
        s.run(sc.dt, n)
        m = s.get_mesh()
        st = s.stats()
    g = -9.8
    assert st["status"] == 0, st
    assert np.all(np.isfinite(m["vx"])) and np.all(np.isfinite(m["D"]))
    assert np.allclose(m["vv"][1], n * sc.dt * g, rtol=1e-4)
    assert np.max(np.abs(m["vv"][[0, 2]])) < 1e-6
    # rigid fall: the element directions keep their rest values
    assert np.max(np.abs(m["D"] - m0["D"])) < 1e-4 * np.max(np.abs(m0["D"]))


def test_cloth_alone_60_substeps():
    """A cloth in a context without particles (the reference's main.cpp:77-84 binds no
    ParticleSystem): same parity bar as the drape test."""
    sc, x, v, faces, fixed = cloth_scene(fixed_corners=True, stretch=True)
    n = 60
    say(f"[cloth alone] vertices {x.shape[1]}, faces {len(faces)}, substeps {n}")
    o = Oracle(sc.params, sc.gmin, sc.gmax, sc.res, prims=sc.prims)
    o.set_mesh(x, faces, v=v, fixed=fixed)
    o.run(sc.dt, n)
    mr = o.get_mesh()
    with Solver() as s:
        s.set_grid(sc.gmin, sc.gmax, sc.res)
        s.set_params(sc.params)
        s.set_levelset(sc.prims)
        s.set_mesh(x, faces, v=v, fixed=fixed)
        s.run(sc.dt, n)
        m = s.get_mesh()
        st = s.stats()
        # a second cloth replaces the first in a context without particles
        s.set_mesh(x, faces, v=v, fixed=fixed)
        s.run(sc.dt, n)
        m2 = s.get_mesh()
    h = 1.0 / sc.res[0]
    dxm = np.max(np.abs(m["vx"] - mr["vx"]))
    say(f"  mesh |dx|/h {dxm / h:.3e}, D rel {rel(m['D'], mr['D']):.3e}")
    assert np.max(np.abs(mr["vx"] - x)) > 10 * h * 1e-3  # it moved
    assert dxm <= 1e-4 * h
    assert rel(m["D"], mr["D"]) < 1e-3
    idx = np.flatnonzero(fixed)
    assert np.array_equal(m["vx"][:, idx], x[:, idx])
    for k in ("vx", "vv", "D"):
        assert np.array_equal(m[k], m2[k]), k
    assert st["status"] == 0, st


def test_cpp_cloth_example(tmp_path):
    """The C++ host API's cloth pattern (LagrangianMesh::ObjMesh, bindConstraints, a solver with no
    ParticleSystem, adaptive solve writing mesh/mesh_N.obj)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "anisotropicelastoplasticity_amd",
                       "cloth_example")
    say("[c++ host api] cloth_example")
    res = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    say(res.stdout.strip())
    assert res.returncode == 0, res.stdout + res.stderr
    nv, nf = 24 * 24, 2 * 23 * 23
    for f in range(2):
        lines = (tmp_path / "mesh" / f"mesh_{f}.obj").read_text().splitlines()
        assert len(lines) == nv + nf
        assert all(l.startswith("v ") for l in lines[:nv]) and all(l.startswith("f ") for l in lines[nv:])
        x, faces = api.read_mesh_obj(str(tmp_path / "mesh" / f"mesh_{f}.obj"))
        assert x.shape == (3, nv) and faces.shape == (nf, 3) and np.all(np.isfinite(x))
    assert not (tmp_path / "mesh" / "mesh_2.obj").exists()
    assert not (tmp_path / "particle" / "particle_0.obj").exists()


def _plane_cloth(plane, n_side=16, edge=1.0 / 40.0, seed=9):
    """A cloth in the plane y = 0.6 or z = 0.6 with the stretching, shearing and bending velocity
    field of cloth_scene, alone in a 32^3 grid with box walls."""
    x, faces = scenes.grid_mesh(n_side, n_side, edge, (0.5 - edge * (n_side - 1) / 2, 0.5 - edge * (n_side - 1) / 2), 0.6)
    if plane == "z":
        x = x[[0, 2, 1]].copy()  # (x, y, z) <- (x, z, y): the sheet lies in z = 0.6
    v = np.zeros_like(x)
    v[0] = (x[0] - 0.5) * 2.0
    inplane = 2 if plane == "y" else 1
    normal = 1 if plane == "y" else 2
    v[inplane] = -(x[inplane] - 0.5) * 1.5
    v[normal] = -0.5 + 0.3 * np.sin(25.0 * x[0]) * np.cos(18.0 * x[inplane])
    prims = [levelset(_lib.LS_BOX, (0.05, 0.05, 0.05, 0.0), (0.95, 0.95, 0.95, 0.0))]
    sc = scenes.small_scene(_lib.SAND, 2, 0, _lib.FRICTION_COULOMB, seed=seed, res=(32, 32, 32), prims=prims, dt=1e-4)
    return sc, x, v.astype(np.float32), faces


def test_cloth_reference_normal_row_z_sheet():
    """mpm_mesh_params.reference_normal_row = 1: the normal-force direction vector is
    (D_rest^T).col(2) as in HybridSolver.cpp:425-427,432. For a sheet in a z = const plane it equals
    D_rest^-T e3 exactly; the device follows the FP64 oracle with the drape test's bar (mesh
    |dx| <= 1e-4 h, D relative L2 <= 1e-3) after 40 substeps."""
    sc, x, v, faces = _plane_cloth("z")
    n = 40
    say(f"[cloth reference normal row] plane z, vertices {x.shape[1]}, substeps {n}")
    res = {}
    for ref_row in (True, False):
        o = Oracle(sc.params, sc.gmin, sc.gmax, sc.res, prims=sc.prims)
        o.set_mesh(x, faces, v=v, reference_normal_row=ref_row)
        o.run(sc.dt, n)
        res[ref_row] = o.get_mesh()
    with Solver() as s:
        s.set_grid(sc.gmin, sc.gmax, sc.res)
        s.set_params(sc.params)
        s.set_levelset(sc.prims)
        s.set_mesh(x, faces, v=v, reference_normal_row=True)
        s.run(sc.dt, n)
        m = s.get_mesh()
        st = s.stats()
    h = 1.0 / sc.res[0]
    mr = res[True]
    dxm = np.max(np.abs(m["vx"] - mr["vx"]))
    d_forms = rel(res[True]["D"], res[False]["D"])
    say(f"  mesh |dx|/h {dxm / h:.3e}, D rel {rel(m['D'], mr['D']):.3e}, oracle forms differ by {d_forms:.3e}")
    assert d_forms == 0.0  # the two forms agree for a z = const sheet
    assert dxm <= 1e-4 * h
    assert rel(m["D"], mr["D"]) < 1e-3
    assert st["status"] == 0, st


def test_cloth_reference_normal_row_y_sheet():
    """reference_normal_row = 1 on a sheet in y = 0.6, where (D_rest^T).col(2) (HybridSolver.cpp:432)
    is not the sheet normal: for rest directions d1, d2 in the x-z plane it is (d1_z, d2_z, 0), so
    the reference form drops the normal and shear stress column of dF3 and takes h times an in-plane
    one. A sand block moving down at 2 m/s lands on the cloth, which compresses the elements' d3
    (r33 < 1), and with normal stiffness 1e7 and shear stiffness 1e6 the two forms part clearly:
    after 40 substeps the FP64 oracle's forms differ by about 3e-4 in D and in |dx|/h, some thirty
    times the FP32 oracle's own distance from the FP64 oracle. The device with the flag set stays
    within twice that FP32 distance (plus 1e-6) of the FP64 oracle with the flag set, and is at
    least five times farther from the oracle without it."""
    prims = [levelset(_lib.LS_BOX, (0.05, 0.05, 0.05, 0.0), (0.95, 0.95, 0.95, 0.0)),
             levelset(_lib.LS_SPHERE, (0.5, 0.42, 0.5, 0.15))]
    sc = scenes.small_scene(_lib.SAND, 2, 0, _lib.FRICTION_COULOMB, seed=5, res=(32, 32, 32), cells=(3, 2, 3),
                            lo=(0.44, 0.61, 0.44), hi=(0.56, 0.6725, 0.56), v0=(0.0, -2.0, 0.0), prims=prims, dt=1e-4)
    edge = 1.0 / 48.0
    n_side = 20
    side = edge * (n_side - 1)
    x, faces = scenes.grid_mesh(n_side, n_side, edge, (0.5 - side / 2, 0.5 - side / 2), 0.6)
    v = np.zeros_like(x)
    k, gamma, n = 1e7, 1e6, 40
    h = 1.0 / sc.res[0]
    say(f"[cloth reference normal row] plane y, particles {sc.n}, vertices {x.shape[1]}, substeps {n}")
    ref = {}
    for ref_row, prec in ((True, 64), (False, 64), (True, 32)):
        o = Oracle(sc.params, sc.gmin, sc.gmax, sc.res, prims=sc.prims, precision=prec)
        o.load_scene(sc)
        o.set_mesh(x, faces, v=v, reference_normal_row=ref_row, k=k, gamma=gamma)
        o.run(sc.dt, n)
        ref[ref_row, prec] = o.get_mesh()
    with Solver() as s:
        sc.load_into(s)
        s.set_mesh(x, faces, v=v, stiffness=k, shear_stiffness=gamma, reference_normal_row=True)
        s.run(sc.dt, n)
        m = s.get_mesh()
        st = s.stats()

    def dist(a, b):
        return np.max(np.abs(a["vx"] - b["vx"])) / h, rel(a["D"], b["D"])

    forms = dist(ref[True, 64], ref[False, 64])
    spread = dist(ref[True, 32], ref[True, 64])
    gpu = dist(m, ref[True, 64])
    gpu_other = dist(m, ref[False, 64])
    say(f"  |dx|/h, D rel: oracle forms {forms[0]:.3e} {forms[1]:.3e}; fp32 oracle {spread[0]:.3e} {spread[1]:.3e}; "
        f"gpu {gpu[0]:.3e} {gpu[1]:.3e}; gpu vs the other form {gpu_other[0]:.3e} {gpu_other[1]:.3e}")
    for i in range(2):
        assert forms[i] > 10.0 * spread[i]  # the flag matters in this scene
        assert gpu[i] <= 2.0 * spread[i] + 1e-6
        assert gpu_other[i] >= 5.0 * gpu[i]
    assert st["status"] == 0, st


def test_cloth_fixed_vertices_at_grid_faces():
    """Fixed vertices one step from the x faces stop the nodes their neighbours' linear indices name
    (HybridSolver.cpp:536-539): the neighbour (-1, j, k) of the low vertex is node (31, j - 1, k) and
    the neighbour (32, j, k) of the high vertex is node (0, j + 1, k). A cloth at rest spans the
    grid in x, so those nodes carry mass; after one P2G and grid update the device stops the same
    active nodes as the FP64 oracle and leaves v = dt g on the others."""
    n_x, n_z = 40, 6
    edge = 0.975 / (n_x - 1)
    x, faces = scenes.grid_mesh(n_x, n_z, edge, (0.01, 0.4), 0.6)
    fixed = np.zeros(x.shape[1], np.uint8)
    lo_v, hi_v = 0 * n_z + 5, (n_x - 1) * n_z + 0  # (x 0.01, z 0.525) and (x 0.985, z 0.4)
    fixed[[lo_v, hi_v]] = 1
    p = scenes.sand_params()
    res = (32, 32, 32)
    dt = 1e-4
    say(f"[cloth fixed at faces] vertices {x.shape[1]}, fixed {np.flatnonzero(fixed).tolist()}")
    o = Oracle(p, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), res)
    o.set_mesh(x, faces, fixed=fixed)
    o.stage(_lib.STAGE_P2G, dt)
    o.stage(_lib.STAGE_GRID, dt)
    g64 = o.get_grid()
    with Solver() as s:
        s.set_grid((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), res)
        s.set_params(p)
        s.set_levelset([])
        s.set_mesh(x, faces, fixed=fixed)
        for st in (_lib.STAGE_SORT, _lib.STAGE_P2G, _lib.STAGE_GRID):
            s.stage(st, dt)
        g = s.get_grid()
        stats = s.stats()
    on = g64["m"] > 0
    assert np.array_equal(g["m"] > 0, on), "active node sets differ"
    stop64 = on & np.all(g64["v"] == 0.0, axis=0)
    stop = on & np.all(g["v"] == 0.0, axis=0) & np.all(g["vbf"] == 0.0, axis=0)

    def node(i, j, k):
        return (k * res[1] + j) * res[0] + i

    # nodes stopped only through the wrap, and active: (31, 18, 17) from the low vertex (stencil
    # j 18..20, k 16..18) and (0, 19, 12) from the high vertex (stencil j 18..20, k 12..14)
    for i, j, k in [(31, 18, 17), (0, 19, 12)]:
        assert on[node(i, j, k)] and stop64[node(i, j, k)], (i, j, k)
    assert np.array_equal(stop, stop64)
    free = on & ~stop64
    assert np.allclose(g["v"][1, free], -9.8 * dt, rtol=1e-5)
    assert stats["status"] == 0, stats
