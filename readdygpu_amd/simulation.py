"""Host-side mirror of the reference's Simulation / CudaSimulation interface.

``Simulation`` drives libCReaDDyGPU.so through its C ABI.  Method names follow
the reference's C++ methods (lib/ReaDDyGPU_Cpp/Simulation.hpp:161-170,
CudaKernels/CudaSimulation.cuh:40-71) in Python spelling; arguments are
numpy arrays in particle-index order.  Every call raises ``RdError`` on a
nonzero status, as the reference's methods return 1 and print.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import hashlib
import json
import os
from typing import Optional

import numpy as np

from . import _lib
from .model import SystemSpec


class RdError(RuntimeError):
    pass


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Simulation:
    """Periodic BD on one GPU (Simulation + CudaSimulation).

    A plain system uses the box of ``spec``.  With ``replicas=(bounds, counts)``
    or ``(bounds, counts, streams)`` the handle holds independent boxes
    (bounds [R,6] as lo,hi per axis, counts [R] particles each, stored
    consecutively in ``xyz``/``types``) that advance together; replica r uses
    noise stream ``streams[r]`` (default r).
    """

    def __init__(self, spec: SystemSpec, xyz: np.ndarray, types: np.ndarray, device: int = 0,
                 rdf: Optional[tuple] = None, initialize: bool = True, replicas: Optional[tuple] = None):
        self.lib = _lib.load()
        self.spec = spec
        self.device = device
        self._tune_tags = {}          # what else shapes a step's cost, for the tuning cache key
        prm = _lib.rd_params()
        for d in range(3):
            prm.lattice_bounds[2 * d] = spec.box_lo[d]
            prm.lattice_bounds[2 * d + 1] = spec.box_hi[d]
        prm.dt = spec.dt
        prm.boltzmann = spec.kB
        prm.temperature = spec.T
        prm.max_cutoff = spec.max_cutoff
        prm.cell_size = spec.cell_size
        prm.seed = spec.seed
        prm.stream = spec.stream
        prm.device = device
        h = C.c_void_p()
        rc = self.lib.rd_create(C.byref(prm), C.byref(h))
        if rc != 0:
            raise RdError(f"rd_create failed ({rc}): no HIP device or bad parameters")
        self.h = h
        nt = spec.n_types
        D = np.array([t.D for t in spec.types], dtype=np.float64)
        rad = np.array([t.radius for t in spec.types], dtype=np.float64)
        radii = np.zeros((nt, nt), dtype=np.float64)
        for i, t in enumerate(spec.types):
            if t.collision_radii is not None:
                radii[i] = np.asarray(t.collision_radii, dtype=np.float64)
        self._check(self.lib.rd_set_particle_types(h, nt, _ptr(D), _ptr(rad), _ptr(radii)))
        for p in spec.pot1:
            q = _lib.rd_pot1()
            q.type, q.subtype, q.force_const = p.kind, p.subtype, p.k
            for d in range(3):
                q.origin[d] = p.origin[d]
                q.normal[d] = p.normal[d]
                q.extension[d] = p.extension[d]
            q.height, q.radius = p.height, p.radius
            q.type_mask = spec.pot1_mask(p)
            self._check(self.lib.rd_add_order_one_potential(h, C.byref(q)))
        for p in spec.pot2:
            q = _lib.rd_pot2()
            q.type, q.subtype, q.force_const, q.length, q.depth = p.kind, p.subtype, p.k, p.length, p.depth
            m = np.ascontiguousarray(spec.pot2_mask(p), dtype=np.uint8)
            self._check(self.lib.rd_add_order_two_potential(h, C.byref(q), _ptr(m)))
        if spec.reaction_radii is not None:
            R = np.ascontiguousarray(spec.reaction_radii, dtype=np.float64)
            self._check(self.lib.rd_set_reaction_radii(h, _ptr(R)))
        if spec.bonds is not None and len(spec.bonds) > 0:
            self.set_bonds(spec.bonds, spec.bond_pot, spec.bond_pots, spec.bond_r0)
        if replicas is None:
            self.set_particles(xyz, types)
        else:
            self.set_replicas(*replicas[:2], xyz, types, streams=replicas[2] if len(replicas) > 2 else None)
        if rdf is not None:
            self.configure_rdf(*rdf)
        if initialize:
            self.initialize()

    # -- plumbing --
    def _check(self, rc: int):
        if rc != 0:
            msg = self.lib.rd_last_error(self.h)
            raise RdError(f"rd_* returned {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.rd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- reference interface --
    def set_particles(self, xyz: np.ndarray, types: np.ndarray):
        """copyPosToDevice + copyTypesToDevice (CudaSimulation.cu:905-970)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        if len(xyz) != len(types):
            raise ValueError("xyz and types differ in length")
        self._check(self.lib.rd_set_particles(self.h, len(types), _ptr(xyz), _ptr(types)))

    def set_bonds(self, bonds: np.ndarray, pot_index, pots, r0=None):
        """Bonded pairs of group potentials (the groups kernel,
        CudaSimulation.cu:1829-1836): bonds [nb,2] global particle indices,
        pot_index [nb] (None: all 0) into ``pots`` (OrderTwoPotential), r0 per
        potential (None or < 0: collision radius sum of the pair's types)."""
        b = np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1, 2)
        nb = len(b)
        pi = (np.zeros(nb, dtype=np.int32) if pot_index is None
              else np.ascontiguousarray(pot_index, dtype=np.int32).reshape(-1))
        if len(pi) != nb:
            raise ValueError("pot_index and bonds differ in length")
        arr = (_lib.rd_pot2 * max(len(pots), 1))()
        for i, p in enumerate(pots):
            arr[i].type, arr[i].subtype, arr[i].force_const = p.kind, p.subtype, p.k
            arr[i].length, arr[i].depth = p.length, p.depth
        r = None
        if r0 is not None:
            r = np.array([-1.0 if v is None else float(v) for v in r0], dtype=np.float64)
            if len(r) != len(pots):
                raise ValueError("r0 needs one value per bond potential")
        self._check(self.lib.rd_set_bonds(self.h, nb, _ptr(b), _ptr(pi), len(pots), arr, _ptr(r)))
        self._tune_tags["bonds"] = nb

    def set_replicas(self, bounds: np.ndarray, counts: np.ndarray, xyz: np.ndarray, types: np.ndarray,
                     streams: Optional[np.ndarray] = None):
        """Independent boxes in one handle (no counterpart in the reference)."""
        b = np.ascontiguousarray(bounds, dtype=np.float64).reshape(-1, 6)
        c = np.ascontiguousarray(counts, dtype=np.int32).reshape(-1)
        if len(b) != len(c):
            raise ValueError("bounds and counts differ in length")
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        if len(xyz) != len(types) or len(types) != int(c.sum()):
            raise ValueError("particle arrays do not match the replica counts")
        st = None if streams is None else np.ascontiguousarray(streams, dtype=np.uint32).reshape(-1)
        if st is not None and len(st) != len(c):
            raise ValueError("streams and counts differ in length")
        self._check(self.lib.rd_set_replicas(self.h, len(c), _ptr(b), _ptr(st), _ptr(c), _ptr(xyz), _ptr(types)))
        self._tune_tags["replicas"] = hashlib.sha1(b.tobytes() + c.tobytes()).hexdigest()

    def initialize(self):
        """CudaSimulation::initialize (CudaSimulation.cu:138-455)."""
        self._check(self.lib.rd_initialize(self.h))

    def simulate(self, nsteps: int = 1):
        """Simulation::simulate(int) (Simulation.cpp:349-355)."""
        self._check(self.lib.rd_simulate(self.h, int(nsteps)))

    def copy_pos_from_device(self):
        """copyPosFromDevice (CudaSimulation.cu:972-984) -> (xyz [N,3] f32, types [N])."""
        n = self.info().n
        xyz = np.empty((n, 3), dtype=np.float32)
        t = np.empty(n, dtype=np.int32)
        self._check(self.lib.rd_copy_pos_from_device(self.h, _ptr(xyz), _ptr(t)))
        return xyz, t

    def cell_ids(self) -> np.ndarray:
        """Reference cell index (CudaSimulation.cu:586-588) per particle."""
        n = self.info().n
        f = np.empty(n, dtype=np.int32)
        self._check(self.lib.rd_cell_ids(self.h, _ptr(f)))
        return f

    def gradients(self):
        """grad U [N,3] (f32) and total energy (FP64 accumulation)."""
        n = self.info().n
        g = np.empty((n, 3), dtype=np.float32)
        e = C.c_double()
        self._check(self.lib.rd_gradients(self.h, _ptr(g), C.byref(e)))
        return g, e.value

    def noise(self, step0: int, nsteps: int) -> np.ndarray:
        """The FP32 N(0,1) array the step uses: [nsteps, 3, N]."""
        n = self.info().n
        out = np.empty((nsteps, 3, n), dtype=np.float32)
        self._check(self.lib.rd_noise_fill(self.h, int(step0), int(nsteps), _ptr(out)))
        return out

    def set_noise_array(self, noise: Optional[np.ndarray]):
        """Take the given noise [nsteps, 3, N] for the next nsteps steps (None: generator)."""
        if noise is None:
            self._check(self.lib.rd_set_noise_array(self.h, None, 0))
            return
        noise = np.ascontiguousarray(noise, dtype=np.float32)
        self._noise_keep = noise
        self._check(self.lib.rd_set_noise_array(self.h, _ptr(noise), noise.shape[0]))

    def configure_rdf(self, r_max: float, n_bins: int):
        self._check(self.lib.rd_rdf_configure(self.h, float(r_max), int(n_bins)))
        nt = self.spec.n_types
        nrep = self.info().n_replicas
        self._rdf_shape = (nt, nt, int(n_bins)) if nrep == 1 else (nrep, nt, nt, int(n_bins))

    def call_rdf_calculation(self):
        """callRDFCalculation (CudaSimulation.cu:505-520)."""
        self._check(self.lib.rd_rdf_calculate(self.h))

    def copy_rdf_matrix(self):
        """copyRDFMatrix (CudaSimulation.cu:457-481): raw ordered-pair counts."""
        c = np.zeros(self._rdf_shape, dtype=np.uint64)
        frames = C.c_long()
        self._check(self.lib.rd_rdf_copy(self.h, _ptr(c), C.byref(frames)))
        return c, frames.value

    def normalize_rdf(self) -> np.ndarray:
        """normalizeRDF (Simulation.cpp:40-62), FP64 with shell and box volume."""
        g = np.zeros(self._rdf_shape, dtype=np.float64)
        self._check(self.lib.rd_rdf_normalize(self.h, _ptr(g)))
        return g

    def _pairs(self, fn) -> np.ndarray:
        # one call with a buffer sized from the previous result; a second one
        # only when the set outgrew it
        key = getattr(fn, "__name__", str(id(fn)))
        cap = max(4096, int(1.25 * getattr(self, "_pair_sizes", {}).get(key, 0)))
        cnt = C.c_long()
        out = np.empty((cap, 2), dtype=np.int64)
        self._check(fn(self.h, _ptr(out), cap, C.byref(cnt)))
        if cnt.value > cap:
            out = np.empty((cnt.value, 2), dtype=np.int64)
            self._check(fn(self.h, _ptr(out), cnt.value, C.byref(cnt)))
        if not hasattr(self, "_pair_sizes"):
            self._pair_sizes = {}
        self._pair_sizes[key] = cnt.value
        return out[:cnt.value]

    def neighbor_pairs(self) -> np.ndarray:
        return self._pairs(self.lib.rd_neighbor_pairs)

    def reaction_candidates(self) -> np.ndarray:
        return self._pairs(self.lib.rd_reaction_candidates)

    def reaction_candidates_noalloc(self, cap: int) -> np.ndarray:
        """rd_reaction_candidates_noalloc: the same set, computed in the
        library's fixed device scratch (no device allocation); at most `cap`
        pairs are returned; the full count is kept in .last_count."""
        cnt = C.c_long()
        out = np.empty((max(cap, 1), 2), dtype=np.int64)
        self._check(self.lib.rd_reaction_candidates_noalloc(self.h, _ptr(out), cap, C.byref(cnt)))
        res = out[:min(cap, cnt.value)]
        self.last_count = cnt.value
        return res

    def checkpoint(self, path: str):
        """Binary snapshot: particles, replica layout, bonds, seed, step counter."""
        self._check(self.lib.rd_checkpoint_write(self.h, path.encode()))

    def restore(self, path: str):
        """Load a snapshot written by a handle with the same model; the run
        continues bit for bit."""
        self._check(self.lib.rd_checkpoint_read(self.h, path.encode()))

    def check_invariants(self) -> int:
        """Number of violated invariants of the cell-sorted state (0: consistent)."""
        v = C.c_long()
        self._check(self.lib.rd_check_invariants(self.h, C.byref(v)))
        return v.value

    def set_step(self, step: int):
        """Set the step (noise) counter."""
        self._check(self.lib.rd_set_step(self.h, int(step)))

    def simulate_async(self, nsteps: int):
        """Enqueue steps without waiting (errors surface at frame_wait)."""
        self._check(self.lib.rd_simulate_async(self.h, int(nsteps)))

    def frame_copy(self):
        """Start copying the current positions/types (index order) to the host."""
        self._check(self.lib.rd_frame_copy(self.h))

    def frame_wait(self, copy: bool = True):
        """Positions [N,3] and types [N] of the last frame_copy; copy=False
        returns views of the pinned host buffers, valid until the next
        frame_copy."""
        px = C.POINTER(C.c_float)()
        pt = C.POINTER(C.c_int)()
        n = C.c_int()
        self._check(self.lib.rd_frame_wait(self.h, C.byref(px), C.byref(pt), C.byref(n)))
        k = n.value
        if k == 0:
            return np.zeros((0, 3), np.float32), np.zeros(0, np.int32)
        xyz = np.ctypeslib.as_array(px, shape=(3 * k,)).reshape(k, 3)
        types = np.ctypeslib.as_array(pt, shape=(k,))
        return (xyz.copy(), types.copy()) if copy else (xyz, types)

    def frame_rewind(self):
        """Back to the state of the last frame_copy (same noise follows)."""
        self._check(self.lib.rd_frame_rewind(self.h))

    def apply_reactions(self, creactions: np.ndarray, creation_replica: Optional[np.ndarray] = None):
        """cReactions protocol (lib/ReaDDyGPU.cpp:303-463); in a handle of
        replicas, entry r that only creates goes to replica creation_replica[r]."""
        cr = np.ascontiguousarray(creactions, dtype=np.float32)
        if creation_replica is None:
            self._check(self.lib.rd_apply_reactions(self.h, _ptr(cr)))
        else:
            rr = np.ascontiguousarray(creation_replica, dtype=np.int32)
            if len(rr) < int(cr[0]):
                raise ValueError("creation_replica needs one entry per reaction")
            self._check(self.lib.rd_apply_reactions_replicas(self.h, _ptr(cr), _ptr(rr)))

    def set_launch_config(self, kernel: int = 0, block: int = 0, ppt: int = 0):
        """Force kernel (0 automatic, 1 per-particle gather with one workgroup
        per group of 64-slot work items, 2 the same gather on a resident grid
        whose workgroups take their share of the work items from LDS tickets), its workgroup size (64, 128,
        256, 512; 0 = 256) and particles per lane (1, 2, 4; 0 = 1; the queue
        form takes one)."""
        c = _lib.rd_launch_config()
        c.kernel = int(kernel)
        c.block = int(block)
        c.ppt = int(ppt)
        self._check(self.lib.rd_set_launch_config(self.h, C.byref(c)))

    def launch_config(self) -> dict:
        c = _lib.rd_launch_config()
        self._check(self.lib.rd_get_launch_config(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in c._fields_}

    def autotune(self, trial_steps: int = 20, cache: Optional[str] = None) -> float:
        """Pick the cell size, force-kernel form, workgroup size and particles
        per lane by timing;
        state is restored.  ``cache`` (default: $RDGPU_TUNE_CACHE) names a JSON
        file of earlier choices keyed by the device and the workload (spec,
        particle count, replica layout, bonds, composite): a hit applies the
        stored choice without timing and returns its recorded ms/step, a miss
        tunes and adds the choice to the file."""
        path = cache if cache is not None else os.environ.get("RDGPU_TUNE_CACHE")
        key = self.tune_key() if path else None
        db = {}
        if path and os.path.exists(path):
            try:
                with open(path) as f:
                    db = json.load(f)
            except (OSError, ValueError):
                db = {}
            hit = db.get(key) if isinstance(db, dict) else None
            if hit:
                self.set_launch_config(kernel=int(hit.get("kernel", 1)), block=int(hit["block"]),
                                       ppt=int(hit["ppt"]))
                self._check(self.lib.rd_set_cell_size(self.h, float(hit["cell_size"])))
                return float(hit["ms_per_step"])
        ms = C.c_double()
        self._check(self.lib.rd_autotune(self.h, int(trial_steps), C.byref(ms)))
        if path:
            lc = self.launch_config()
            db = db if isinstance(db, dict) else {}
            db[key] = {"cell_size": self.info().cell_size, "kernel": lc["used_kernel"], "block": lc["used_block"],
                       "ppt": lc["used_ppt"], "ms_per_step": ms.value}
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(db, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        return ms.value

    def set_cell_size(self, cell: float):
        """Lattice cell edge of the running system (<= 0: automatic); positions
        and the step counter are kept."""
        self._check(self.lib.rd_set_cell_size(self.h, float(cell)))

    def tune_key(self) -> str:
        """Key of this workload in the tuning cache: device name and a digest of
        the spec, the particle count and the replica/bond/composite layout."""
        buf = C.create_string_buffer(256)
        self._check(self.lib.rd_device_name(int(self.device), buf, 256))
        h = hashlib.sha1()

        def feed(v):
            if dataclasses.is_dataclass(v) and not isinstance(v, type):
                for f in dataclasses.fields(v):
                    h.update(f.name.encode())
                    feed(getattr(v, f.name))
            elif isinstance(v, np.ndarray):
                h.update(str(v.dtype).encode() + str(v.shape).encode())
                h.update(np.ascontiguousarray(v).tobytes())
            elif isinstance(v, (list, tuple)):
                h.update(b"[")
                for x in v:
                    feed(x)
                h.update(b"]")
            elif isinstance(v, dict):
                for k in sorted(v):
                    h.update(str(k).encode())
                    feed(v[k])
            else:
                h.update(repr(v).encode())
        feed(self.spec)
        feed(int(self.info().n))
        feed(self._tune_tags)
        return f"{buf.value.decode()}|{h.hexdigest()}"

    # -- composite external potential and surrogate (no reference counterpart) --
    def set_composite(self, pieces, types=None, combine: str = "nearest"):
        """Composite of attractive DISK / CYLINDER pieces
        (model.OrderOnePotential), combined by nearest piece or, with
        combine="sum", summed as orderOne sums every order-one potential;
        applies to `types` (default: all)."""
        arr = (_lib.rd_pot1 * max(len(pieces), 1))()
        for i, p in enumerate(pieces):
            q = arr[i]
            q.type, q.subtype, q.force_const = p.kind, p.subtype, p.k
            for d in range(3):
                q.origin[d] = p.origin[d]
                q.normal[d] = p.normal[d]
                q.extension[d] = p.extension[d]
            q.height, q.radius = p.height, p.radius
            q.type_mask = 0xFFFFFFFF
        mask = 0xFFFFFFFF if types is None else sum(1 << int(t) for t in types)
        self._check(self.lib.rd_set_composite(self.h, len(pieces), C.cast(arr, C.c_void_p), mask))
        if combine not in ("nearest", "sum"):
            raise ValueError("combine is 'nearest' or 'sum'")
        self._check(self.lib.rd_set_composite_combine(self.h, 1 if combine == "sum" else 0))
        self._tune_tags["composite"] = [len(pieces), mask, combine]

    def set_external_mode(self, surrogate: bool):
        self._check(self.lib.rd_set_external_mode(self.h, 1 if surrogate else 0))
        self._tune_tags["surrogate"] = bool(surrogate)

    def composite_eval(self, xyz):
        """Exact composite on the GPU at positions [m,3] -> (grad [m,3], energy [m])."""
        x = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        g = np.zeros_like(x)
        e = np.zeros(len(x), dtype=np.float32)
        self._check(self.lib.rd_composite_eval(self.h, len(x), _ptr(x), _ptr(g), _ptr(e)))
        return g, e

    def surrogate_init(self, seed: int = 1):
        self._check(self.lib.rd_surrogate_init(self.h, int(seed)))

    def surrogate_weights(self) -> np.ndarray:
        w = np.zeros(8771, dtype=np.float32)
        self._check(self.lib.rd_surrogate_get_weights(self.h, _ptr(w)))
        return w

    def set_surrogate_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != 8771:
            raise ValueError("the surrogate has 8771 parameters")
        self._check(self.lib.rd_surrogate_set_weights(self.h, _ptr(w)))

    def surrogate_norm(self):
        lo = np.zeros(3)
        hi = np.zeros(3)
        gs = C.c_double()
        self._check(self.lib.rd_surrogate_get_norm(self.h, _ptr(lo), _ptr(hi), C.byref(gs)))
        return lo, hi, gs.value

    def set_surrogate_norm(self, lo, hi, gscale):
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        self._check(self.lib.rd_surrogate_set_norm(self.h, _ptr(lo), _ptr(hi), float(gscale)))

    def surrogate_make_data(self, nsamples: int, seed: int = 1) -> float:
        gs = C.c_double()
        self._check(self.lib.rd_surrogate_make_data(self.h, int(nsamples), int(seed), C.byref(gs)))
        return gs.value

    def surrogate_train(self, steps: int, batch: int = 8192, lr: float = 3e-3, lr_min: float = 1e-5,
                        seed: int = 1) -> np.ndarray:
        tp = _lib.rd_train_params()
        tp.batch, tp.steps, tp.lr, tp.lr_min = int(batch), int(steps), lr, lr_min
        tp.beta1, tp.beta2, tp.eps, tp.seed = 0.9, 0.999, 1e-8, int(seed)
        curve = np.zeros(int(steps), dtype=np.float32)
        self._check(self.lib.rd_surrogate_train(self.h, C.byref(tp), _ptr(curve)))
        return curve

    def surrogate_loss_grad(self, xn, t):
        x = np.ascontiguousarray(xn, dtype=np.float32).reshape(-1, 3)
        tt = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 3)
        g = np.zeros(8771, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.rd_surrogate_loss_grad(self.h, len(x), _ptr(x), _ptr(tt), C.byref(loss), _ptr(g)))
        return loss.value, g

    def surrogate_predict(self, xn):
        x = np.ascontiguousarray(xn, dtype=np.float32).reshape(-1, 3)
        y = np.zeros_like(x)
        self._check(self.lib.rd_surrogate_predict(self.h, len(x), _ptr(x), _ptr(y)))
        return y

    def surrogate_validate(self, nsamples: int = 200_000, seed: int = 99) -> float:
        r = C.c_double()
        self._check(self.lib.rd_surrogate_validate(self.h, int(nsamples), int(seed), C.byref(r)))
        return r.value

    def info(self) -> _lib.rd_info:
        i = _lib.rd_info()
        self._check(self.lib.rd_get_info(self.h, C.byref(i)))
        return i
