"""Learned warm start of the Newton solve: the network, its graphs, training data from the solver, training.

The network (drape-2_amd/csrc/warmstart.hpp) predicts from the state each step's velocity residual beyond free
flight, d = dv - h (g + f_ext / m) (dv the step's velocity change); attached to a `HipSolver` it seeds Newton with
the inertial target plus h d, u = h v + h^2 (g + f_ext / m) + h d, instead of the reference's u = h v
(cpu_solver.cpp:244), and the solver keeps the plain guess whenever the prediction's energy is higher. Training runs on the GPU through the
network's own forward/backward kernels; this module only moves arrays and drives the loop.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _ffi
from ._ffi import load_library, ptr

IN, EDGE, HIDDEN, ROUNDS, OUT, OUT_USED = 16, 8, 64, 3, 16, 3
VEL_SCALE, DV_SCALE, LEN_SCALE, PHI_SCALE = 1.0, 0.5, 0.01, 0.02


def _lib():
    return load_library()


def _check(code: int) -> None:
    _ffi.check(code, _lib().drape_last_error)


def param_layout() -> dict:
    """name -> (offset, shape) of each tensor in the flat parameter vector (row-major [in][out])."""
    out, o = {}, 0

    def take(name, shape):
        nonlocal o
        out[name] = (o, shape)
        o += int(np.prod(shape))

    take("We", (IN, HIDDEN))
    take("be", (HIDDEN,))
    for r in range(ROUNDS):
        take(f"Wab{r}", (HIDDEN, 2 * HIDDEN))
        take(f"Wc{r}", (EDGE, HIDDEN))
        take(f"bm{r}", (HIDDEN,))
        take(f"Wu{r}", (2 * HIDDEN, HIDDEN))
        take(f"bu{r}", (HIDDEN,))
    take("Wd", (HIDDEN, OUT))
    take("bd", (OUT,))
    out["_total"] = (o, ())
    return out


def scene_graph(scene):
    """Triangle-edge graph of a scene in particle order (sorted unique neighbours) and rest positions:
    (row_start uint32 [n+1], nbr uint32, rest2 float64 [n, 2])."""
    tris, rest, off = [], [], 0
    for g in scene.garments:
        tris.append(np.asarray(g.mesh.triangles, np.int64) + off)
        rest.append(np.asarray(g.mesh.rest, np.float64))
        off += g.mesh.n
    n = off
    t = np.concatenate(tris) if tris else np.zeros((0, 3), np.int64)
    a = np.concatenate([t[:, 0], t[:, 1], t[:, 2], t[:, 1], t[:, 2], t[:, 0]])
    b = np.concatenate([t[:, 1], t[:, 2], t[:, 0], t[:, 0], t[:, 1], t[:, 2]])
    key = np.unique(a * n + b)
    src, dst = key // n, key % n
    rs = np.zeros(n + 1, np.uint32)
    np.add.at(rs, src + 1, 1)
    rs = np.cumsum(rs).astype(np.uint32)
    return rs, dst.astype(np.uint32), np.concatenate(rest) if rest else np.zeros((0, 2))


class MeshGraph:
    """A mesh edge graph on the device."""

    def __init__(self, row_start, nbr, rest2):
        self.row_start = np.ascontiguousarray(row_start, np.uint32)
        self.nbr = np.ascontiguousarray(nbr, np.uint32)
        self.rest2 = np.ascontiguousarray(rest2, np.float64).reshape(-1, 2)
        self.n = len(self.row_start) - 1
        self.handle = C.c_void_p()
        _check(_lib().drape_graph_create(C.byref(self.handle), self.n, ptr(self.row_start, C.c_uint32),
                                         ptr(self.nbr if len(self.nbr) else np.zeros(1, np.uint32), C.c_uint32),
                                         ptr(self.rest2, C.c_double)))

    @classmethod
    def of_scene(cls, scene) -> "MeshGraph":
        return cls(*scene_graph(scene))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _ffi._lib is not None:
            _ffi._lib.drape_graph_destroy(h)
            self.handle = None


def _frame_arrays(frame: dict):
    return [np.ascontiguousarray(frame[k], np.float32) for k in ("x", "v", "dv_prev", "phi_normal")]


class WarmStartNet:
    """The warm-start network: weights, Adam state and working buffers on the device."""

    def __init__(self, seed: int = 0):
        self.handle = C.c_void_p()
        _check(_lib().drape_net_create(C.byref(self.handle), int(seed)))
        n = C.c_size_t()
        _check(_lib().drape_net_param_count(self.handle, C.byref(n)))
        self.n_params = n.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _ffi._lib is not None:
            _ffi._lib.drape_net_destroy(h)
            self.handle = None

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        _check(_lib().drape_net_get_params(self.handle, ptr(out, C.c_float), self.n_params))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        _check(_lib().drape_net_set_params(self.handle, ptr(p, C.c_float), len(p)))

    def reset_optimizer(self) -> None:
        _check(_lib().drape_net_reset_optimizer(self.handle))

    def predict(self, graph: MeshGraph, frame: dict) -> np.ndarray:
        x, v, d, pn = _frame_arrays(frame)
        out = np.empty((graph.n, 3), np.float32)
        _check(_lib().drape_net_predict(self.handle, graph.handle, ptr(x, C.c_float), ptr(v, C.c_float),
                                        ptr(d, C.c_float), ptr(pn, C.c_float), ptr(out, C.c_float)))
        return out

    def step(self, graph: MeshGraph, frame: dict, target: np.ndarray, lr: float, want_grad: bool = False):
        """Loss (and gradient when asked) on one frame; lr > 0 takes an Adam step."""
        x, v, d, pn = _frame_arrays(frame)
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_double()
        grad = np.empty(self.n_params, np.float32) if want_grad else None
        _check(_lib().drape_net_train_step(self.handle, graph.handle, ptr(x, C.c_float), ptr(v, C.c_float),
                                           ptr(d, C.c_float), ptr(pn, C.c_float), ptr(t, C.c_float), float(lr),
                                           C.byref(loss), ptr(grad, C.c_float)))
        return (loss.value, grad) if want_grad else loss.value

    def save(self, path: str) -> None:
        np.save(path, self.params())

    def load(self, path: str) -> None:
        self.set_params(np.load(path, allow_pickle=False))


@dataclasses.dataclass
class Sample:
    graph: MeshGraph
    frame: dict
    target: np.ndarray


_replicas: dict = {}


def replicate(graph: MeshGraph, k: int) -> MeshGraph:
    """k disjoint copies of a graph (block-diagonal), cached: the graph of k frames of one mesh stacked."""
    key = (id(graph), k)
    if key not in _replicas:
        n, e = graph.n, len(graph.nbr)
        rs = np.concatenate([graph.row_start[:-1].astype(np.int64) + i * e for i in range(k)] + [[k * e]])
        nbr = np.concatenate([graph.nbr.astype(np.int64) + i * n for i in range(k)])
        _replicas[key] = (graph, MeshGraph(rs, nbr, np.tile(graph.rest2, (k, 1))))
    return _replicas[key][1]


def stack(samples: list) -> Sample:
    """Samples of one mesh as one sample over the replicated graph."""
    g = replicate(samples[0].graph, len(samples))
    frame = {key: np.concatenate([s.frame[key] for s in samples]) for key in samples[0].frame}
    return Sample(g, frame, np.concatenate([s.target for s in samples]))


def residual_target(scene, dv: np.ndarray, h: float) -> np.ndarray:
    """The network's target for a realised velocity change dv: its residual beyond free flight, dv - h g for a
    free particle of a settling garment (the guess is then the inertial target plus h times the residual), zero
    for a pinned one. Scenes with external forces are not used for training."""
    if scene.forces:
        raise ValueError("training scenes carry no external forces")
    out = dv.astype(np.float32) - np.float32(h) * np.asarray(scene.settings["gravity"], np.float32)
    if scene.pins:
        out[[p for p, _ in scene.pins]] = 0.0
    return out


def collect(scene, steps: int, dt: float = 1 / 60, iterations: int = 5, every: int = 1, solver=None,
            graph: MeshGraph | None = None) -> list:
    """Solver-generated training pairs: the inputs of the state before a step and the velocity residual the step
    realised beyond free flight (residual_target). The plain (previous-velocity) guess is used; every `every`-th
    step is kept. Pass `graph` to share one graph between scenes of the same mesh (so that their frames can be
    stacked)."""
    from .solver import HipSolver
    s = solver or HipSolver()
    if solver is None:
        s.build(scene)
    graph = graph or MeshGraph.of_scene(scene)
    out = []
    for k in range(steps):
        frame = s.frame_inputs() if k % every == 0 else None
        s.step(dt, 1, iterations)
        if frame is not None:
            out.append(Sample(graph, frame, residual_target(scene, s.frame_inputs()["dv_prev"], dt)))
    return out


def train(net: WarmStartNet, samples: list, epochs: int, lr: float = 1e-3, seed: int = 0, log=None,
          batch_vertices: int = 0, lr_final: float | None = None) -> list:
    """Adam over the samples in a seeded random order per epoch; returns the mean loss of each epoch.
    With batch_vertices > 0, samples of the same mesh are stacked into batches of about that many vertices.
    The learning rate decays geometrically to lr_final (default: lr) over the epochs."""
    rng = np.random.default_rng(seed)
    history = []
    for e in range(epochs):
        rate = lr if lr_final is None or epochs == 1 else lr * (lr_final / lr) ** (e / (epochs - 1))
        order = list(rng.permutation(len(samples)))
        batches = []
        if batch_vertices > 0:
            by_graph: dict = {}
            for i in order:
                by_graph.setdefault(id(samples[i].graph), []).append(samples[i])
            for group in by_graph.values():
                k = max(1, batch_vertices // group[0].graph.n)
                batches += [group[j:j + k] for j in range(0, len(group), k)]
            batches = [batches[i] for i in rng.permutation(len(batches))]
        else:
            batches = [[samples[i]] for i in order]
        tot = 0.0
        for b in batches:
            s = b[0] if len(b) == 1 else stack(b)
            tot += net.step(s.graph, s.frame, s.target, rate) * len(b)
        history.append(tot / max(len(samples), 1))
        if log:
            log(f"epoch {e}: lr {rate:.2e} loss {history[-1]:.5f}")
    return history
