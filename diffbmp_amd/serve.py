"""Serving many independent fit jobs on one GPU.

A fit job is the reference's optimisation loop for one image
(``pydiffbmp/core/renderer/simple_tile_renderer.py:1043-1140``: render,
MSE to the target, backward, Adam, clamp) with its own canvas, primitive
count, template set and target.  Here jobs arrive over time, are admitted
while their device memory fits an HBM budget, and every admitted job advances
one Adam step per *batched step*: one launch of each stage (binning, raster
forward, background composite + MSE + raster backward, Adam) covers all of
them (``dbmp_batch_step``).  A batched step can be captured into a hipGraph
once per job set and replayed.

* ``JobSpec`` / ``FitState``   a job's shape and its device buffers;
* ``job_bytes``                the device bytes a job holds while admitted;
* ``HbmBudget``                admission accounting;
* ``BatchEngine``              the native batch: job set, steps, pair capacity;
* ``Server``                   arrivals, admission, completion, latencies.
"""
from __future__ import annotations

import ctypes
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native
from ._native import Bitmaps, Job, check

_QUAD_BYTES = {torch.float32: 16, torch.float16: 8, torch.bfloat16: 8}
_RGBA_BYTES = {torch.float32: 64, torch.float16: 16, torch.bfloat16: 16}


@dataclass
class JobSpec:
    """Shape of one fit job.  ``template_ids`` index the shared template
    library (T = len(template_ids))."""
    height: int
    width: int
    num_primitives: int
    template_ids: tuple
    steps: int = 200
    seed: int = 0
    lr: float = 1e-2
    c_blend: float = 0.5
    eps_T: float = 1e-6
    alpha_upper_bound: float = 1.0
    ls_range: tuple = (0.005, 0.1)    # log-uniform radius scale r = e^ls * W / 2
    arrival: float = 0.0              # seconds from the start of the run
    name: str = ""
    template_size: int = 256
    library: str = "stress"           # template_library kind
    clamp: bool = True                # the reference's clamp of x, y, r after each step


def job_bytes(spec: JobSpec, store_dtype=torch.float16) -> int:
    """Device bytes of one admitted job (the tensors ``FitState`` allocates)."""
    H, W, N, T = spec.height, spec.width, spec.num_primitives, len(spec.template_ids)
    Ht = Wt = spec.template_size
    es = torch.tensor([], dtype=store_dtype).element_size()
    b = N * 8 * 4 * 3 + N * 8 * es + N * 4          # master, moments, store, sel
    rec = _RGBA_BYTES[store_dtype] if spec.c_blend > 0 else _QUAD_BYTES[store_dtype]
    b += T * Ht * Wt * (es + 3 * es) + T * (Ht + 1) * Wt * rec + T * 16   # bitmaps, maps, records, extents
    b += H * W * (3 + 3 + 1 + 1 + 1) * 4             # target, colour, alpha, T, last
    b += spec.steps * 4                              # loss history
    return int(b)


class HbmBudget:
    """Admission accounting against a byte budget (no allocation itself)."""

    def __init__(self, capacity_bytes: int):
        self.capacity = int(capacity_bytes)
        self.used = 0

    def fits(self, nbytes: int) -> bool:
        return self.used + int(nbytes) <= self.capacity

    def reserve(self, nbytes: int) -> bool:
        if not self.fits(nbytes):
            return False
        self.used += int(nbytes)
        return True

    def release(self, nbytes: int) -> None:
        self.used -= int(nbytes)
        if self.used < 0:
            raise RuntimeError("HBM budget released more than it reserved")


# --------------------------------------------------------------------------
# synthetic jobs (template library of the stress workload, smooth targets)
# --------------------------------------------------------------------------
_LIBRARY: dict = {}


def template_library(size: int = 256, count: int = 16, kind: str = "stress"):
    """Alpha templates (count, size, size) and per-template colour maps
    (count, size, size, 3), fp32 numpy.  ``kind`` "stress": the stress
    workload's RGBA bitmaps; "hard": binary centre squares of inset
    size/8 .. 3 size/8 (hard-edged, content short of the bitmap edge) with the
    stress workload's colour maps."""
    key = (size, count, kind)
    if key not in _LIBRARY:
        from . import scenes as S
        sc = S.make_scene(3, H=64, W=64, N=1, tsz=size, with_target=False)
        tm, cm = sc.templates[:count].copy(), sc.colors_orig_tmpl[:count].copy()
        if kind == "hard":
            tm = np.stack([S.hard_square(size, size // 8 + (i * size) // (4 * max(count, 1)))
                           for i in range(count)])
        elif kind != "stress":
            raise ValueError(f"unknown template library {kind!r}")
        _LIBRARY[key] = (tm, cm)
    return _LIBRARY[key]


_DEVICE_LIBRARY: dict = {}


def device_library(size: int, kind: str, dtype, device):
    """``template_library`` on the device in ``dtype``, uploaded once per
    (size, kind, dtype, device); jobs select their bitmaps from it."""
    key = (size, kind, dtype, str(device))
    if key not in _DEVICE_LIBRARY:
        tm, cm = template_library(size, kind=kind)
        _DEVICE_LIBRARY[key] = (torch.from_numpy(tm).to(device).to(dtype).contiguous(),
                                torch.from_numpy(cm).to(device).to(dtype).contiguous())
    return _DEVICE_LIBRARY[key]


def smooth_target(H: int, W: int, seed: int, device, terms: int = 32) -> torch.Tensor:
    """``clamp(0.5 + sum_i (a_i/8) sin(2 pi (kx_i x/W + ky_i y/H) + phi_i), 0, 1)``
    per channel (a ~ U[0,1], k ~ U[0.5,8], phi ~ U[0,2 pi)), evaluated on the GPU.

    Each term separates, sin(p(x) + q(y)) = sin p cos q + cos p sin q, so a
    channel is 0.5 + U V with U (H, 2 terms) the row factors and V (2 terms, W)
    the column factors: one batched product for the three channels instead of
    a few hundred elementwise kernels."""
    rng = np.random.default_rng(seed)
    coef = np.empty((4, 3, terms), np.float64)   # a, kx, ky, phi per channel, in draw order
    for ch in range(3):
        coef[0, ch] = rng.uniform(0, 1, terms)
        coef[1, ch] = rng.uniform(0.5, 8, terms)
        coef[2, ch] = rng.uniform(0.5, 8, terms)
        coef[3, ch] = rng.uniform(0, 2 * math.pi, terms)
    a, kx, ky, ph = torch.from_numpy(coef.astype(np.float32)).to(device).unbind(0)   # (3, terms)
    y = torch.arange(H, device=device, dtype=torch.float32) / H
    x = torch.arange(W, device=device, dtype=torch.float32) / W
    q = (2 * math.pi) * ky[:, None, :] * y[None, :, None] + ph[:, None, :]          # (3, H, terms)
    p = (2 * math.pi) * kx[:, :, None] * x[None, None, :]                           # (3, terms, W)
    a8 = (a / 8.0)[:, None, :]
    U = torch.cat((a8 * torch.cos(q), a8 * torch.sin(q)), dim=2)                    # (3, H, 2 terms)
    V = torch.cat((torch.sin(p), torch.cos(p)), dim=1)                              # (3, 2 terms, W)
    f = torch.bmm(U, V).add_(0.5).clamp_(0, 1)                                      # (3, H, W)
    return f.permute(1, 2, 0).contiguous()


def random_spec(rng: np.random.Generator, steps: int = 200, library_size: int = 16,
                canvases=(256, 512, 1024, 2048), n_range=(512, 32768)) -> JobSpec:
    """A serving job: canvas side uniform over ``canvases``, primitive count
    log-uniform over ``n_range``, 1..16 template types."""
    side = int(rng.choice(canvases))
    N = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
    T = int(rng.integers(1, library_size + 1))
    ids = tuple(int(i) for i in rng.choice(library_size, T, replace=False))
    return JobSpec(side, side, N, ids, steps=steps, seed=int(rng.integers(1 << 30)))


def prepare_bitmaps(batch: "BatchEngine", templates, colors_orig, blend: bool):
    """Extents and texel records (RGBA records with colour maps) of a template
    set, in the engine's storage type: ``(ext, quads)``."""
    dev, sdt = batch.device, batch.store_dtype
    T, Ht, Wt = templates.shape
    ext = torch.empty((T, 4), dtype=torch.float32, device=dev)
    bm = Bitmaps(templates.data_ptr(), colors_orig.data_ptr() if blend else None, T, Ht, Wt,
                 _native._DT[sdt], 2 if blend else 0, 0.0, -1, 0, -1)
    nbytes = int(batch._L.dbmp_batch_bitmap_bytes(batch._b, ctypes.byref(bm)))
    if nbytes < 0:
        raise ValueError(f"bad template set {(T, Ht, Wt)}")
    quads = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        check(batch._L.dbmp_batch_prepare_bitmaps(batch._b, ctypes.byref(bm), _native._ptr(ext),
                                                 _native._ptr(quads), _native._stream(dev)))
    return ext, quads


class FitState:
    """Device buffers of one admitted job."""

    def __init__(self, spec: JobSpec, batch: "BatchEngine"):
        dev, sdt = batch.device, batch.store_dtype
        self.spec = spec
        H, W, N = spec.height, spec.width, spec.num_primitives
        rng = np.random.default_rng(spec.seed)
        T = len(spec.template_ids)
        self.blend = spec.c_blend > 0
        if batch.share_bitmaps:
            # the engine's one copy of the library's texel records; the job's
            # primitives index the library directly
            self.templates, self.colors_orig, self.ext, self.quads = batch.shared_bitmaps(
                spec.template_size, spec.library, self.blend)
            sel_map = np.asarray(spec.template_ids, np.int32)
        else:
            lib_t, lib_c = device_library(spec.template_size, spec.library, sdt, dev)
            ids = torch.tensor(list(spec.template_ids), dtype=torch.long, device=dev)
            self.templates = lib_t.index_select(0, ids)
            self.colors_orig = lib_c.index_select(0, ids)
            self.ext, self.quads = prepare_bitmaps(batch, self.templates, self.colors_orig, self.blend)
            sel_map = np.arange(T, dtype=np.int32)
        lo, hi = spec.ls_range
        x = rng.uniform(0, W, N)
        y = rng.uniform(0, H, N)
        r = np.exp(rng.uniform(math.log(lo), math.log(hi), N)) * W / 2.0
        th = rng.uniform(0, 2 * math.pi, N)
        v = rng.standard_normal(N)
        u = rng.uniform(0.001, 0.999, (N, 3))
        c = np.log(u / (1 - u))
        p = np.concatenate([np.stack([x, y, r, th, v], 1), c], 1).astype(np.float32)
        self.params = torch.from_numpy(p).to(dev).contiguous()
        self.exp_avg = torch.zeros_like(self.params)
        self.exp_avg_sq = torch.zeros_like(self.params)
        self.store = self.params if sdt == torch.float32 else self.params.to(sdt)
        self.sel = torch.from_numpy(sel_map[(np.arange(N) % T)[::-1]].astype(np.int32)).to(dev)
        self.target = smooth_target(H, W, spec.seed + 1, dev)
        self.color = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
        self.alpha = torch.empty((H, W), dtype=torch.float32, device=dev)
        self.T = torch.empty((H, W), dtype=torch.float32, device=dev)
        self.last = torch.empty((H, W), dtype=torch.int32, device=dev)
        self.loss = torch.zeros(max(spec.steps, 1), dtype=torch.float32, device=dev)
        self.steps_done = 0
        self.nbytes = job_bytes(spec, sdt)
        # timestamps (host clock, seconds)
        self.t_admit = None
        self.t_done = None
        self.done_event = None

    def c_job(self, exact_blend: bool = False) -> Job:
        s = self.spec
        T, Ht, Wt = self.templates.shape
        j = Job()
        j.height, j.width, j.num_primitives = s.height, s.width, s.num_primitives
        j.num_templates, j.tmpl_height, j.tmpl_width = T, Ht, Wt
        j.quads, j.ext = self.quads.data_ptr(), self.ext.data_ptr()
        j.quads_with_color = int(self.blend)
        j.c_blend = float(s.c_blend)
        j.blend_grad_exact = int(exact_blend)
        j.bmp_sel = self.sel.data_ptr()
        j.params, j.exp_avg, j.exp_avg_sq = (self.params.data_ptr(), self.exp_avg.data_ptr(),
                                             self.exp_avg_sq.data_ptr())
        j.store = self.store.data_ptr()
        j.target = self.target.data_ptr()
        j.background = None
        j.background_value = 1.0
        j.out_color, j.out_alpha = self.color.data_ptr(), self.alpha.data_ptr()
        j.out_T, j.out_last = self.T.data_ptr(), self.last.data_ptr()
        j.loss_history, j.loss_capacity = self.loss.data_ptr(), self.loss.numel()
        for i in range(5):
            j.lr[i] = float(s.lr)
        j.alpha_upper_bound = float(s.alpha_upper_bound)
        j.eps_T = float(s.eps_T)
        j.clamp = int(s.clamp)
        j.steps_done = self.steps_done
        return j

    def unpack(self) -> dict:
        """Parameters as the reference's five tensors (fp32 master copies)."""
        p = self.params
        return dict(means2D=p[:, 0:2].clone(), radii=p[:, 2].clone(), rotations=p[:, 3].clone(),
                    opacities=p[:, 4].clone(), colors=p[:, 5:8].clone())


class BatchEngine:
    """The native batch (``dbmp_batch_*``) and its current job set."""

    def __init__(self, store_dtype=torch.float16, betas=(0.9, 0.999), eps=1e-8,
                 use_graph: bool = True, pairs_per_primitive: float = 64.0, device=None,
                 exact_blend: bool = False, share_bitmaps: bool = True):
        """``share_bitmaps``: jobs whose bitmaps come from the same library
        share one copy of its texel records (the same samples from a
        sixteenth or less of the distinct memory) instead of each holding
        the records of its own subset."""
        self._L = _native.lib()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.store_dtype = store_dtype
        self.use_graph = bool(use_graph)
        self.exact_blend = bool(exact_blend)
        self.pairs_per_primitive = float(pairs_per_primitive)
        b = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self._L.dbmp_batch_create(ctypes.byref(b), _native._DT[store_dtype],
                                            float(betas[0]), float(betas[1]), float(eps)))
        self._b = b
        self.jobs: list[FitState] = []
        self._dirty = True
        self.pair_capacity = 0
        self._array = None
        self.set_count = 0
        self.share_bitmaps = bool(share_bitmaps)
        self._shared: dict = {}

    def shared_bitmaps(self, size: int, kind: str, blend: bool):
        """The library ``(templates, colour maps, extents, records)`` on this
        engine's device and storage type, built on first use."""
        key = (size, kind, bool(blend))
        if key not in self._shared:
            tm, cm = device_library(size, kind, self.store_dtype, self.device)
            self._shared[key] = (tm, cm, *prepare_bitmaps(self, tm, cm, blend))
        return self._shared[key]

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            try:
                self._L.dbmp_batch_destroy(b)
            except Exception:
                pass
            self._b = None

    def add(self, job: FitState) -> None:
        self.jobs.append(job)
        self._dirty = True

    def remove(self, job: FitState) -> None:
        self.jobs.remove(job)
        self._dirty = True

    def _upload(self) -> None:
        n = sum(j.spec.num_primitives for j in self.jobs)
        want = int(self.pairs_per_primitive * max(n, 1))
        self.pair_capacity = min(max(self.pair_capacity, want), 0x7FFFFFFF)
        arr = (Job * max(len(self.jobs), 1))()
        for i, j in enumerate(self.jobs):
            arr[i] = j.c_job(self.exact_blend)
        self._array = arr
        check(self._L.dbmp_batch_set_jobs(self._b, arr, len(self.jobs), self.pair_capacity,
                                          _native._stream(self.device)))
        self._dirty = False
        self.set_count += 1

    def step(self) -> None:
        """One Adam step of every job in the set, queued on the current stream."""
        if not self.jobs:
            return
        if self._dirty:
            self._upload()
        check(self._L.dbmp_batch_step(self._b, int(self.use_graph), _native._stream(self.device)))
        for j in self.jobs:
            j.steps_done += 1

    def gradients(self, job: FitState) -> torch.Tensor:
        """(N,8) gradient rows of ``job`` from the most recent step (x, y, r,
        theta, opacity logit, 3 colour logits), before its Adam update."""
        out = torch.empty((job.spec.num_primitives, 8), dtype=torch.float32, device=self.device)
        check(self._L.dbmp_batch_gradients(self._b, self.jobs.index(job), _native._ptr(out),
                                           _native._stream(self.device)))
        return out

    def pair_count(self) -> int:
        """Pairs of the most recent completed step (host-mapped mirror)."""
        v = ctypes.c_int64()
        check(self._L.dbmp_batch_pair_count(self._b, ctypes.byref(v)))
        return int(v.value)

    def grow_if_needed(self) -> bool:
        """After a completed step: raise the pair capacity if that step
        overflowed it (the step was exact, only slower)."""
        p = self.pair_count()
        if p > self.pair_capacity:
            n = sum(j.spec.num_primitives for j in self.jobs)
            self.pairs_per_primitive = max(self.pairs_per_primitive, 1.3 * p / max(n, 1))
            self.pair_capacity = int(1.3 * p)
            self._dirty = True
            return True
        return False


@dataclass
class ServeResult:
    spec: JobSpec
    latency: float          # completion - arrival (s)
    queue: float            # admission - arrival (s)
    final_loss: float
    first_loss: float


@dataclass
class ServeStats:
    results: list = field(default_factory=list)
    wall: float = 0.0
    steps: int = 0
    max_active: int = 0

    def summary(self) -> dict:
        lat = np.array([r.latency for r in self.results]) if self.results else np.zeros(1)
        return dict(jobs=len(self.results), wall_s=self.wall,
                    jobs_per_s=len(self.results) / self.wall if self.wall > 0 else 0.0,
                    latency_p50_s=float(np.percentile(lat, 50)),
                    latency_p99_s=float(np.percentile(lat, 99)),
                    batched_steps=self.steps, max_active=self.max_active)


class GpuMarker:
    """A point in the engine's stream (a recorded torch event)."""

    def __init__(self, device):
        self.ev = torch.cuda.Event()
        self.ev.record(torch.cuda.current_stream(device))

    def query(self) -> bool:
        return self.ev.query()

    def synchronize(self) -> None:
        self.ev.synchronize()


class Server:
    """Admits jobs in arrival order while they fit the HBM budget and the
    slot limit, steps every admitted job once per batched step, and retires
    a job when its last step has completed on the device.

    ``make_state(spec)`` builds an admitted job (default ``FitState``) and
    ``make_marker()`` marks the stream after a step (default ``GpuMarker``);
    both can be replaced to run the admission logic without a GPU."""

    def __init__(self, engine, budget: HbmBudget, max_jobs: int = 32,
                 max_in_flight: int = 2, clock=time.perf_counter, make_state=None,
                 make_marker=None):
        self.engine = engine
        self.budget = budget
        self.max_jobs = int(max_jobs)
        self.max_in_flight = int(max_in_flight)
        self.clock = clock
        self.make_state = make_state or (lambda spec: FitState(spec, engine))
        self.make_marker = make_marker or (lambda: GpuMarker(engine.device))

    def run(self, specs: list[JobSpec], keep_params: bool = False) -> ServeStats:
        eng = self.engine
        queue = deque(sorted(specs, key=lambda s: s.arrival))
        for s in queue:
            if job_bytes(s, eng.store_dtype) > self.budget.capacity:
                raise ValueError(f"job {s.name or s.seed} needs {job_bytes(s, eng.store_dtype)} B, "
                                 f"more than the whole budget")
        finishing: list[FitState] = []
        in_flight: deque = deque()
        stats = ServeStats()
        t0 = self.clock()
        self.params_out = {}
        while queue or eng.jobs or finishing:
            now = self.clock() - t0
            # admission: FIFO while the head fits
            while (queue and queue[0].arrival <= now and len(eng.jobs) < self.max_jobs
                   and self.budget.fits(job_bytes(queue[0], eng.store_dtype))):
                spec = queue.popleft()
                self.budget.reserve(job_bytes(spec, eng.store_dtype))
                js = self.make_state(spec)
                js.t_admit = self.clock() - t0
                eng.add(js)
            if not eng.jobs and not finishing:
                if queue:
                    time.sleep(max(0.0, min(queue[0].arrival - now, 0.01)))
                continue
            if eng.jobs:
                eng.step()
                stats.steps += 1
                stats.max_active = max(stats.max_active, len(eng.jobs))
                ev = self.make_marker()
                in_flight.append(ev)
                for js in [j for j in eng.jobs if j.steps_done >= j.spec.steps]:
                    js.done_event = ev
                    eng.remove(js)
                    finishing.append(js)
                while len(in_flight) > self.max_in_flight:
                    in_flight.popleft().synchronize()
                    eng.grow_if_needed()
            else:
                while in_flight:
                    in_flight.popleft().synchronize()
            for js in [j for j in finishing if j.done_event.query()]:
                js.t_done = self.clock() - t0
                finishing.remove(js)
                loss = js.loss.cpu() if hasattr(js.loss, "cpu") else np.asarray(js.loss)
                stats.results.append(ServeResult(js.spec, js.t_done - js.spec.arrival,
                                                 js.t_admit - js.spec.arrival,
                                                 float(loss[js.spec.steps - 1]), float(loss[0])))
                if keep_params:
                    self.params_out[js.spec.name or js.spec.seed] = (js.unpack(), loss)
                self.budget.release(js.nbytes)
        while in_flight:
            in_flight.popleft().synchronize()
        stats.wall = self.clock() - t0
        return stats


def poisson_arrivals(rate: float, n: int, rng: np.random.Generator) -> np.ndarray:
    """Arrival times (s) of a Poisson process of ``rate`` jobs/s."""
    return np.cumsum(rng.exponential(1.0 / rate, n))
