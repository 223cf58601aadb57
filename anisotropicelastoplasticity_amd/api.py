"""Python host wrapper over the C ABI (include/mpm_c_abi.h), used by the tests and the benchmark.

`Solver` owns one device context (one HIP stream + device buffers) and mirrors the C entry points
one to one; arrays are numpy float32 in the planar layout of the C ABI (x has shape (3, n), B and F
have shape (9, n) row-major, scalars have shape (n,)).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import LevelSetPrim, MeshParams, Params, SceneDesc, Stats, Tuning


class MPMError(RuntimeError):
    pass


def _fp(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _planar(a, rows, n, name):
    if a is None:
        return None
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(rows, n) if rows > 1
                               else np.asarray(a, dtype=np.float32).reshape(n))
    return arr


class Solver:
    """One simulation context on one GPU (HybridSolver of HybridSolver.h:27-95, over the C ABI)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.mpm_create(ctypes.byref(h), int(device))
        if rc != 0:
            raise MPMError(f"mpm_create failed ({rc}): {self.lib.mpm_last_error(None).decode()}")
        self.h = h
        self.n = 0
        self.res = None

    # ---------------------------------------------------------------- plumbing
    def _check(self, rc, what):
        if rc != 0:
            raise MPMError(f"{what} failed ({rc}): {self.lib.mpm_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.mpm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------------- setup
    def set_grid(self, min_bound, max_bound, res):
        mn = (ctypes.c_double * 3)(*[float(v) for v in min_bound])
        mx = (ctypes.c_double * 3)(*[float(v) for v in max_bound])
        rs = (ctypes.c_int32 * 3)(*[int(v) for v in res])
        self._check(self.lib.mpm_set_grid(self.h, mn, mx, rs), "mpm_set_grid")
        self.res = tuple(int(v) for v in res)

    def set_params(self, params):
        self._check(self.lib.mpm_set_params(self.h, ctypes.byref(params)), "mpm_set_params")

    def set_levelset(self, prims):
        arr = (LevelSetPrim * max(1, len(prims)))()
        for i, p in enumerate(prims):
            arr[i] = p
        self._check(self.lib.mpm_set_levelset(self.h, arr, len(prims)), "mpm_set_levelset")

    def set_particles(self, x, mass, v=None, B=None, F=None, plastic=None, volume=None, phi=None):
        x = np.asarray(x, dtype=np.float32)
        n = x.shape[-1]
        arrs = dict(x=_planar(x, 3, n, "x"), v=_planar(v, 3, n, "v"), B=_planar(B, 9, n, "B"),
                    F=_planar(F, 9, n, "F"), plastic=_planar(plastic, 1, n, "plastic"),
                    mass=_planar(mass, 1, n, "mass"), volume=_planar(volume, 1, n, "volume"),
                    phi=_planar(phi, 1, n, "phi"))
        self._keep = arrs
        rc = self.lib.mpm_set_particles(self.h, n, _fp(arrs["x"]), _fp(arrs["v"]), _fp(arrs["B"]), _fp(arrs["F"]),
                                        _fp(arrs["plastic"]), _fp(arrs["mass"]), _fp(arrs["volume"]),
                                        _fp(arrs["phi"]))
        self._check(rc, "mpm_set_particles")
        self.n = n

    def set_scenes(self, scenes):
        """Several independent scenes in this context (mpm_set_scenes); `scenes` is a list of
        scenes.Scene. Their particles are then set concatenated in the same order (load_batch)."""
        descs = (SceneDesc * len(scenes))()
        for k, sc in enumerate(scenes):
            d = descs[k]
            for a in range(3):
                d.min_bound[a] = float(sc.gmin[a])
                d.max_bound[a] = float(sc.gmax[a])
                d.res[a] = int(sc.res[a])
            d.n_particles = int(sc.n)
            d.params = sc.params
            if len(sc.prims) > _lib.MAX_LEVELSETS:
                raise MPMError("at most 16 level-set primitives per scene")
            d.n_levelsets = len(sc.prims)
            for i, pr in enumerate(sc.prims):
                d.levelsets[i] = pr
        self._check(self.lib.mpm_set_scenes(self.h, len(scenes), descs), "mpm_set_scenes")
        self.res = None
        self.scene_counts = [int(sc.n) for sc in scenes]
        self.scene_nodes = [int(np.prod(sc.res)) for sc in scenes]

    def load_batch(self, scenes):
        """set_scenes + set_particles with every scene's particles concatenated in order."""
        self.set_scenes(scenes)
        keys = ["x", "v", "B", "F", "plastic", "mass", "volume", "phi"]
        cat = {}
        for k in keys:
            parts = [sc.particles.get(k) for sc in scenes]
            if all(p is None for p in parts):
                cat[k] = None
                continue
            filled = []
            for sc, p in zip(scenes, parts):
                if p is None:  # a scene without the field: its default
                    n = sc.n
                    p = {"v": np.zeros((3, n)), "B": np.zeros((9, n)), "phi": np.zeros(n),
                         "F": np.tile(np.eye(3).reshape(9, 1), (1, n)),
                         "plastic": np.full(n, 1.0 if sc.params.material == _lib.SNOW else 0.0)}[k]
                filled.append(np.asarray(p, np.float32))
            cat[k] = np.concatenate(filled, axis=-1)
        self.set_particles(cat["x"], cat["mass"], v=cat["v"], B=cat["B"], F=cat["F"], plastic=cat["plastic"],
                           volume=cat["volume"], phi=cat["phi"])

    def set_mesh(self, x, faces, v=None, fixed=None, youngs_modulus=200.0, poisson_ratio=0.3, shear_stiffness=1e3,
                 stiffness=1e4, friction_angle=20.0, density=2e3, thickness=0.04, reference_normal_row=False):
        """Attach a triangle cloth (mpm_set_mesh) after set_particles, or alone in a context without
        particles (the reference's main.cpp:77-84): x (3, nv), faces (nf, 3)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        nv = x.shape[1]
        f = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1, 3))
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float32)
        fx = None if fixed is None else np.ascontiguousarray(np.asarray(fixed, dtype=np.uint8))
        mp = MeshParams(youngs_modulus, poisson_ratio, shear_stiffness, stiffness, friction_angle, density, thickness,
                        int(reference_normal_row))
        self._keep_mesh = (x, f, vv, fx)
        rc = self.lib.mpm_set_mesh(self.h, nv, _fp(x), _fp(vv), f.shape[0],
                                   f.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   None if fx is None else fx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                   ctypes.byref(mp))
        self._check(rc, "mpm_set_mesh")
        self.nv, self.nf = nv, f.shape[0]

    def get_mesh(self):
        vx = np.empty((3, self.nv), np.float32)
        vv = np.empty((3, self.nv), np.float32)
        D = np.empty((9, self.nf), np.float32)
        self._check(self.lib.mpm_get_mesh(self.h, _fp(vx), _fp(vv), _fp(D)), "mpm_get_mesh")
        return dict(vx=vx, vv=vv, D=D)

    # ---------------------------------------------------------------- stepping
    def estimate_volumes(self):
        self._check(self.lib.mpm_estimate_volumes(self.h), "mpm_estimate_volumes")

    def substep(self, dt):
        self._check(self.lib.mpm_substep(self.h, float(dt)), "mpm_substep")

    def run(self, dt, nsteps):
        self._check(self.lib.mpm_run(self.h, float(dt), int(nsteps)), "mpm_run")

    def solve(self, cfl, maxt, on_frame=None):
        """Adaptive time stepping of HybridSolver::solve (mpm_solve). `on_frame(frame, t)` runs after
        the substep that completes each 1/60 s frame; exceptions it raises stop nothing but are
        re-raised after mpm_solve returns."""
        errors = []

        def _cb(_ctx, frame, t, _user):
            if on_frame is None or errors:
                return
            try:
                on_frame(int(frame), float(t))
            except BaseException as e:  # noqa: BLE001 - re-raised below
                errors.append(e)

        cb = _lib.FRAME_CALLBACK(_cb)
        rc = self.lib.mpm_solve(self.h, float(cfl), float(maxt), ctypes.cast(cb, ctypes.c_void_p), None)
        self._check(rc, "mpm_solve")
        if errors:
            raise errors[0]

    def autotune(self, dt, reps=5):
        t = Tuning()
        self._check(self.lib.mpm_autotune(self.h, float(dt), int(reps), ctypes.byref(t)), "mpm_autotune")
        return {name: getattr(t, name) for name, _ in Tuning._fields_}

    def tuning(self):
        t = Tuning()
        self._check(self.lib.mpm_get_tuning(self.h, ctypes.byref(t)), "mpm_get_tuning")
        return {name: getattr(t, name) for name, _ in Tuning._fields_}

    def set_tuning(self, p2g_ranks, p2g_blocks, grid_blocks, sort_mode=_lib.SORT_COUNTING, p2g_pass=_lib.P2G_ROWS):
        t = Tuning(int(p2g_ranks), int(p2g_blocks), int(grid_blocks), 0.0, 0.0, int(sort_mode), 0.0, int(p2g_pass))
        self._check(self.lib.mpm_set_tuning(self.h, ctypes.byref(t)), "mpm_set_tuning")

    def clock(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        frames, subs = ctypes.c_int32(), ctypes.c_int64()
        self._check(self.lib.mpm_get_clock(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(frames),
                                           ctypes.byref(subs)), "mpm_get_clock")
        return dict(t=t.value, dt=dt.value, frames=frames.value, substeps=subs.value)

    def stage(self, stage, dt):
        self._check(self.lib.mpm_stage(self.h, int(stage), float(dt)), "mpm_stage")

    def synchronize(self):
        self._check(self.lib.mpm_synchronize(self.h), "mpm_synchronize")

    # ---------------------------------------------------------------- readback
    def get_particles(self):
        n = self.n
        out = dict(x=np.empty((3, n), np.float32), v=np.empty((3, n), np.float32), B=np.empty((9, n), np.float32),
                   F=np.empty((9, n), np.float32), plastic=np.empty(n, np.float32), volume=np.empty(n, np.float32))
        rc = self.lib.mpm_get_particles(self.h, _fp(out["x"]), _fp(out["v"]), _fp(out["B"]), _fp(out["F"]),
                                        _fp(out["plastic"]), _fp(out["volume"]))
        self._check(rc, "mpm_get_particles")
        return out

    def get_stress(self):
        K = np.empty((6, self.n), np.float32)
        self._check(self.lib.mpm_get_stress(self.h, _fp(K)), "mpm_get_stress")
        return K

    def set_stress(self, K):
        K = np.ascontiguousarray(K, dtype=np.float32).reshape(6, self.n)
        self._check(self.lib.mpm_set_stress(self.h, _fp(K)), "mpm_set_stress")

    # ---------------------------------------------------------------- frames and checkpoints
    def write_obj(self, path):
        """Particle positions as an OBJ point list, one `v x y z` line per particle in caller order
        (the frame files of HybridSolver.cpp:996-1005)."""
        write_obj(path, self.get_particles()["x"])

    def write_mesh_obj(self, path):
        """The cloth as the reference's mesh frame file (HybridSolver.cpp:1006-1030): `v x y z` per
        vertex, then `f a b c` per face with 1-based indices."""
        write_mesh_obj(path, self.get_mesh()["vx"], self._keep_mesh[1])

    def save_checkpoint(self, path):
        """Resumable snapshot (numpy .npz, no pickles): particle state in caller order, masses,
        per-particle friction angles, the stress state and the substep count."""
        st = self.get_particles()
        keep = self._keep
        data = dict(x=st["x"], v=st["v"], B=st["B"], F=st["F"], plastic=st["plastic"], volume=st["volume"],
                    mass=keep["mass"], substeps=np.int64(self.stats()["substeps"]))
        if keep.get("phi") is not None:
            data["phi"] = keep["phi"]
        try:
            data["stress"] = self.get_stress()
        except MPMError:
            pass  # stress at the predicted F keeps no stress state
        np.savez(path, **data)

    def load_checkpoint(self, path):
        """Restore a snapshot written by save_checkpoint into a context with the same grid and
        parameters; the following substeps repeat the original run bit for bit."""
        with np.load(path, allow_pickle=False) as z:
            self.set_particles(z["x"], z["mass"], v=z["v"], B=z["B"], F=z["F"], plastic=z["plastic"],
                               volume=z["volume"], phi=z["phi"] if "phi" in z.files else None)
            if "stress" in z.files:
                self.set_stress(z["stress"])

    def get_sorted(self, bins=True):
        """Sorted cell keys, caller ids in sorted order and (single scene) the cell bin offsets."""
        n = self.n
        keys = np.empty(n, np.uint32)
        ids = np.empty(n, np.int32)
        b = np.empty(int(np.prod(self.res)) + 1, np.int32) if bins else None
        rc = self.lib.mpm_get_sorted(self.h, keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                     ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     None if b is None else b.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        self._check(rc, "mpm_get_sorted")
        return keys, ids, b

    def get_grid(self, m=True, mom=True, v=True, vbf=True):
        """Dense grid arrays (mpm_get_grid); a batched context returns its scenes' nodes concatenated
        scene by scene. Planes not asked for are not copied (None)."""
        ng = int(np.prod(self.res)) if self.res is not None else int(sum(self.scene_nodes))
        out = dict(m=np.empty(ng, np.float32) if m else None, mom=np.empty((3, ng), np.float32) if mom else None,
                   v=np.empty((3, ng), np.float32) if v else None, vbf=np.empty((3, ng), np.float32) if vbf else None)
        self._check(self.lib.mpm_get_grid(self.h, *[_fp(out[k]) for k in ("m", "mom", "v", "vbf")]), "mpm_get_grid")
        return out

    def stats(self):
        s = Stats()
        self._check(self.lib.mpm_get_stats(self.h, ctypes.byref(s)), "mpm_get_stats")
        return {name: getattr(s, name) for name, _ in Stats._fields_}

    def clear_status(self):
        self._check(self.lib.mpm_clear_status(self.h), "mpm_clear_status")

    # ---------------------------------------------------------------- learned return-map surrogate
    def surrogate_nparam(self):
        return int(self.lib.mpm_surrogate_nparam())

    def surrogate_init(self, seed):
        self._check(self.lib.mpm_surrogate_init(self.h, ctypes.c_uint64(int(seed))), "mpm_surrogate_init")

    def surrogate_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.mpm_surrogate_set_weights(self.h, _fp(w), int(w.size)), "mpm_surrogate_set_weights")

    def surrogate_get_weights(self):
        w = np.empty(self.surrogate_nparam(), np.float32)
        self._check(self.lib.mpm_surrogate_get_weights(self.h, _fp(w), int(w.size)), "mpm_surrogate_get_weights")
        return w

    def surrogate_predict(self, inputs):
        inp = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 5)
        H = np.empty((inp.shape[0], 3), np.float32)
        self._check(self.lib.mpm_surrogate_predict(self.h, _fp(inp), _fp(H), inp.shape[0]), "mpm_surrogate_predict")
        return H

    def surrogate_gradient(self, inputs, H):
        inp = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 5)
        Ht = np.ascontiguousarray(H, dtype=np.float32).reshape(-1, 3)
        g = np.empty(self.surrogate_nparam(), np.float32)
        loss = ctypes.c_float()
        rc = self.lib.mpm_surrogate_gradient(self.h, _fp(inp), _fp(Ht), inp.shape[0], _fp(g), ctypes.byref(loss))
        self._check(rc, "mpm_surrogate_gradient")
        return g, loss.value

    def surrogate_train(self, inputs, H, epochs, batch, lr, seed=0):
        inp = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 5)
        Ht = np.ascontiguousarray(H, dtype=np.float32).reshape(-1, 3)
        losses = np.empty(int(epochs), np.float32)
        rc = self.lib.mpm_surrogate_train(self.h, _fp(inp), _fp(Ht), inp.shape[0], int(epochs), int(batch), float(lr),
                                          ctypes.c_uint64(int(seed)), _fp(losses))
        self._check(rc, "mpm_surrogate_train")
        return losses

    def set_pair_logging(self, capacity):
        self._check(self.lib.mpm_set_pair_logging(self.h, int(capacity)), "mpm_set_pair_logging")

    def get_pairs(self, max_pairs=None):
        if max_pairs is None:
            max_pairs = self.stats()["pairs_logged"]
        out = np.empty((max(int(max_pairs), 1), 8), np.float32)
        cnt = ctypes.c_int64()
        self._check(self.lib.mpm_get_pairs(self.h, _fp(out), int(max_pairs), ctypes.byref(cnt)), "mpm_get_pairs")
        return out[:cnt.value]

    def set_surrogate_log(self, capacity):
        """Record the gate decision of up to `capacity` surrogate candidates per substep."""
        self._check(self.lib.mpm_set_surrogate_log(self.h, int(capacity)), "mpm_set_surrogate_log")

    def get_surrogate_log(self, max_records=1 << 24):
        """The last surrogate substep's gate records: dict of slot (int32), eps (n, 3), phi, kmu,
        H (n, 3) and accepted (bool)."""
        cnt = ctypes.c_int64(0)
        self._check(self.lib.mpm_get_surrogate_log(self.h, None, int(max_records), ctypes.byref(cnt)),
                    "mpm_get_surrogate_log")
        m = int(cnt.value)
        rec = np.empty((max(m, 1), 10), np.float32)
        self._check(self.lib.mpm_get_surrogate_log(self.h, _fp(rec), m, ctypes.byref(cnt)), "mpm_get_surrogate_log")
        rec = rec[:m]
        return dict(slot=rec[:, 0].copy().view(np.int32), eps=rec[:, 1:4], phi=rec[:, 4], kmu=rec[:, 5], H=rec[:, 6:9],
                    accepted=rec[:, 9] > 0.5)

    def set_surrogate(self, enable, gate_tol=1e-5):
        self._check(self.lib.mpm_set_surrogate(self.h, int(bool(enable)), float(gate_tol)), "mpm_set_surrogate")

    def bench_copy(self, nbytes, reps):
        ms = ctypes.c_float()
        self._check(self.lib.mpm_bench_copy(self.h, int(nbytes), int(reps), ctypes.byref(ms)), "mpm_bench_copy")
        return ms.value

    def time_substeps(self, dt, nsteps):
        total = ctypes.c_float()
        stages = (ctypes.c_float * _lib.NUM_STAGES)()
        rc = self.lib.mpm_time_substeps(self.h, float(dt), int(nsteps), ctypes.byref(total), stages)
        self._check(rc, "mpm_time_substeps")
        return total.value, list(stages)

    def time_kernels(self, dt, nsteps):
        """(total ms, per-stage ms, per-stage main-kernel ms) summed over nsteps eager substeps
        (mpm_time_kernels): the G2P entry of the last list is k_g2p alone."""
        total = ctypes.c_float()
        stages = (ctypes.c_float * _lib.NUM_STAGES)()
        kern = (ctypes.c_float * _lib.NUM_STAGES)()
        rc = self.lib.mpm_time_kernels(self.h, float(dt), int(nsteps), ctypes.byref(total), stages, kern)
        self._check(rc, "mpm_time_kernels")
        return total.value, list(stages), list(kern)

    def time_surrogate(self, dt, nsteps):
        """(surrogate kernel ms, candidates) summed over nsteps eager substeps with the surrogate
        enabled (mpm_time_surrogate); the kernel evaluates the network, gates and updates the state."""
        f, c = ctypes.c_float(), ctypes.c_int64()
        rc = self.lib.mpm_time_surrogate(self.h, float(dt), int(nsteps), ctypes.byref(f), ctypes.byref(c))
        self._check(rc, "mpm_time_surrogate")
        return f.value, c.value


def write_obj(path, x):
    """`v x y z` per particle, as the reference's frame writer (HybridSolver.cpp:996-1005)."""
    x = np.asarray(x)
    with open(path, "w") as f:
        for i in range(x.shape[1]):
            f.write("v %.6g %.6g %.6g\n" % (x[0, i], x[1, i], x[2, i]))


def generate_block(seed, lo, hi, cells, ppa):
    """Seeded stratified-jitter particle block (mpm_generate_block); returns x with shape (3, n)."""
    lib = _lib.load()
    n = int(cells[0]) * int(cells[1]) * int(cells[2]) * int(ppa) ** 3
    x = np.empty((3, n), np.float32)
    lo_ = (ctypes.c_double * 3)(*[float(v) for v in lo])
    hi_ = (ctypes.c_double * 3)(*[float(v) for v in hi])
    c_ = (ctypes.c_int32 * 3)(*[int(v) for v in cells])
    rc = lib.mpm_generate_block(ctypes.c_uint64(int(seed)), lo_, hi_, c_, int(ppa), n, _fp(x))
    if rc != 0:
        raise MPMError(f"mpm_generate_block failed ({rc})")
    return x


def generate_shape(seed, shape, a, n):
    """Seeded rejection sampler of the reference's shape factories (mpm_generate_shape): shape is
    _lib.SHAPE_BALL (a = cx, cy, cz, r), SHAPE_BLOCK (a = min xyz, max xyz, hole radius) or
    SHAPE_CYLINDER (a = base xyz, r, height, axis). Returns x with shape (3, n)."""
    lib = _lib.load()
    n = int(n)
    x = np.empty((3, n), np.float32)
    a = [float(v) for v in a]
    a_ = (ctypes.c_double * 8)(*(a + [0.0] * (8 - len(a))))
    rc = lib.mpm_generate_shape(ctypes.c_uint64(int(seed)), int(shape), a_, n, _fp(x))
    if rc != 0:
        raise MPMError(f"mpm_generate_shape failed ({rc})")
    return x


def levelset(kind, a=(0, 0, 0, 0), b=(0, 0, 0, 0)):
    p = LevelSetPrim()
    p.type = int(kind)
    for i in range(4):
        p.a[i] = float(a[i])
        p.b[i] = float(b[i])
    return p


def write_mesh_obj(path, x, faces):
    """`v x y z` per vertex, then `f a b c` (1-based) per face (HybridSolver.cpp:1006-1030)."""
    x = np.asarray(x)
    faces = np.asarray(faces).reshape(-1, 3)
    with open(path, "w") as f:
        for i in range(x.shape[1]):
            f.write("v %.6g %.6g %.6g\n" % (x[0, i], x[1, i], x[2, i]))
        for a, b, c in faces:
            f.write("f %d %d %d\n" % (a + 1, b + 1, c + 1))


def read_mesh_obj(path):
    """LagrangianMesh::ObjMesh's reader (LagrangianMesh.cpp:198-280): `v` and `f` lines, 1-based
    indices (`a/b/c` keeps a; negative indices count from the end), polygons split into fans.
    Returns x (3, nv) float32 and faces (nf, 3) int32."""
    pos, faces = [], []
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                pos.append([float(t) for t in tok[1:4]])
            elif tok[0] == "f":
                idx = []
                for t in tok[1:]:
                    k = int(t.split("/")[0])
                    if k == 0:
                        raise MPMError("%s:%d: face index 0" % (path, lineno))
                    idx.append(k - 1 if k > 0 else len(pos) + k)
                if len(idx) < 3:
                    raise MPMError("%s:%d: face with < 3 corners" % (path, lineno))
                for c in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[c], idx[c + 1]])
    x = np.asarray(pos, np.float32).reshape(-1, 3).T.copy()
    fa = np.asarray(faces, np.int32).reshape(-1, 3)
    if fa.size and (fa.min() < 0 or fa.max() >= x.shape[1]):
        raise MPMError("%s: face index out of range" % path)
    return x, fa
