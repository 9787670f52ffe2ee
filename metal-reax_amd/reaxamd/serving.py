"""Batched and served MD on one GPU: independent systems of uneven size advanced concurrently.

A `Server` owns `slots` worker threads; each worker owns one engine context (its own HIP stream and device buffers) and runs
one request at a time: box, atoms, charge-equilibration setup, NVT (or NVE) MD for the requested steps, then the final
energies, forces and charges. A dispatcher admits queued requests in arrival order as soon as a slot is free and the
request's device memory (its atom count times the bytes per atom the engines report) fits in the budget, so requests that
arrive while others run start at once and their kernels overlap on the device through the workers' streams.

Every engine is deterministic (fixed-order reductions, no atomics on floating point) and a worker resets the charge history
before each request, so a request's result does not depend on what runs beside it: alone or in the mix, the answer is the
same bit for bit (tests/test_gpu_serving.py).

Reference: the reference has no batching or serving (capabilities.cpp:112); the workloads are those of the reference's
benchmark plan (512 systems, N log-uniform in [100, 3000], half water, half CH4:O2 = 1:2 gas).
"""
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import TERMS, Engine, systems


@dataclass
class Request:
    rid: int
    system: object                  # systems.System
    steps: int
    t_target: float = 300.0
    dt: float = 0.25
    t_damp: float = 25.0
    qeq_tol: float = 1e-6
    seed: int = 0
    arrival: float = 0.0            # seconds after the server's start (for Poisson traces)
    # filled in by the server
    start: float = 0.0
    end: float = 0.0
    result: dict = field(default_factory=dict)

    @property
    def atom_steps(self):
        return self.system.n * self.steps


# device bytes per atom of one engine context at the capacities it allocates for condensed C/H/O systems (skin list, bond and
# H-bond candidate lists, bond slots, charge matrix, per-atom state); the server refines it from the engines' own capacities
DEFAULT_BYTES_PER_ATOM = 40_000


def run_one(engine, ffield_text, req, precision):
    """One request on one engine: the plain single-system path (the same calls a user makes)."""
    w = req.system
    e = engine
    e.set_box(w.lo, w.hi)
    e.set_atoms(w.types, w.tags, w.x)
    e.qeq_reset_history()
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    v = systems.maxwell_boltzmann(w.masses(), req.t_target, seed=req.seed)
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    e.md_setup(v, masses, dt=req.dt, t_target=req.t_target, t_damp=req.t_damp, skin=2.0, qeq_every=1, qeq_imax=200,
               qeq_tol=req.qeq_tol)
    thermo = e.md_run(req.steps)
    r = e.compute()
    x, vel, q = e.state()
    return {"energy": r["energy"], "grad": r["grad"], "charges": q, "positions": x, "thermo": thermo}


def run_packed(engine, ffield_text, requests):
    """Requests packed in one context (rxa_set_systems): one launch of each kernel advances them all. They must share the time
    step and the step count; thermostat targets and damping may differ. Returns one result dict per request (as run_one)."""
    if len({r.steps for r in requests}) != 1 or len({r.dt for r in requests}) != 1 or len({r.qeq_tol for r in requests}) != 1:
        raise ValueError("packed requests must share steps, dt and the charge-solve tolerance")
    e = engine
    e.set_systems([r.system for r in requests])
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    v = np.concatenate([systems.maxwell_boltzmann(r.system.masses(), r.t_target, seed=r.seed) for r in requests])
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    r0 = requests[0]
    e.md_setup(v, masses, dt=r0.dt, t_target=r0.t_target, t_damp=r0.t_damp, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=r0.qeq_tol)
    if any(r.t_target != r0.t_target or r.t_damp != r0.t_damp for r in requests):
        e.md_system_targets([r.t_target for r in requests], [r.t_damp for r in requests])
    thermo = e.md_run(r0.steps)
    g = e.compute()["grad"]
    x, vel, q = e.state()
    per = e.system_results()
    out, a = [], 0
    for k, r in enumerate(requests):
        b = a + r.system.n
        out.append({"energy": dict(zip(TERMS, per["energy"][k].tolist())), "grad": g[a:b], "charges": q[a:b], "positions": x[a:b],
                    "thermo": {"ke": float(per["ke"][k]), "temp": float(per["temp"][k]), "iters_s": int(per["iters_s"][k]),
                               "iters_t": int(per["iters_t"][k]), "pack": thermo}})
        a = b
    return out


def run_pack_chunks(engine, ffield_text, requests, t0=None):
    """Requests of any step counts packed in one context: the pack advances in chunks that end where some request's step
    count is reached, and that request's result is taken there (its later steps are not used). Splitting a run changes no
    state and a packed system's results do not depend on the others, so each result is the one the request gets alone.
    Sets start/end (seconds after t0) on the requests."""
    if len({r.dt for r in requests}) != 1 or len({r.qeq_tol for r in requests}) != 1:
        raise ValueError("packed requests must share dt and the charge-solve tolerance")
    now = (lambda: time.perf_counter() - t0) if t0 is not None else (lambda: 0.0)
    e = engine
    start = now()
    for r in requests:
        r.start = start
    e.set_systems([r.system for r in requests])
    e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
    v = np.concatenate([systems.maxwell_boltzmann(r.system.masses(), r.t_target, seed=r.seed) for r in requests])
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    r0 = requests[0]
    e.md_setup(v, masses, dt=r0.dt, t_target=r0.t_target, t_damp=r0.t_damp, skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=r0.qeq_tol)
    if any(r.t_target != r0.t_target or r.t_damp != r0.t_damp for r in requests):
        e.md_system_targets([r.t_target for r in requests], [r.t_damp for r in requests])
    offs = np.concatenate([[0], np.cumsum([r.system.n for r in requests])])
    done = 0
    for target in sorted({r.steps for r in requests}):
        thermo = e.md_run(target - done)
        done = target
        g = e.compute()["grad"]
        x, vel, q = e.state()
        per = e.system_results()
        t = now()
        for k, r in enumerate(requests):
            if r.steps != target:
                continue
            a, b = offs[k], offs[k + 1]
            r.result = {"energy": dict(zip(TERMS, per["energy"][k].tolist())), "grad": g[a:b].copy(), "charges": q[a:b].copy(),
                        "positions": x[a:b].copy(),
                        "thermo": {"ke": float(per["ke"][k]), "temp": float(per["temp"][k]), "iters_s": int(per["iters_s"][k]),
                                   "iters_t": int(per["iters_t"][k]), "pack": thermo}}
            r.end = t
    return requests


class _Moved:
    """A system at its current positions (the continuing member of a re-pack)."""

    def __init__(self, system, x):
        self.x = np.asarray(x, dtype=np.float64)
        self.types, self.tags, self.lo, self.hi = system.types, system.tags, system.lo, system.hi
        self.periodic = getattr(system, "periodic", (True, True, True))
        self.name = system.name

    @property
    def n(self):
        return len(self.types)


def _result(per, g, x, q, k, a, b, thermo):
    return {"energy": dict(zip(TERMS, per["energy"][k].tolist())), "grad": g[a:b].copy(), "charges": q[a:b].copy(),
            "positions": x[a:b].copy(),
            "thermo": {"ke": float(per["ke"][k]), "temp": float(per["temp"][k]), "iters_s": int(per["iters_s"][k]),
                       "iters_t": int(per["iters_t"][k]), "pack": thermo}}


def run_continuous(engine, ffield_text, admit, chunk=50, t0=None, repack_waste=0.25):
    """Continuous batching in one context: the pack advances in chunks of at most `chunk` steps (a chunk also ends where a
    member reaches its step count, whose result is taken there); at a chunk boundary `admit(capacity_atoms, block)` hands over
    newly admitted requests (a list; None once no more will come; block=True when the pack is empty), and the pack is rebuilt
    with them and its unfinished members, which continue from their state (positions, velocities, charges, charge history and
    thermostat chain: rxa_md_state / rxa_md_resume). A finished member rides along until the next re-pack (its later steps are
    not used) unless finished members hold more than `repack_waste` of the pack's atoms. All members share dt and the
    charge-solve tolerance. A continuing member's result equals that of its run without re-packs up to the order of sums (a
    re-pack rebuilds its lists); a pack that does not change keeps the bitwise isolation of packed runs."""
    now = (lambda: time.perf_counter() - t0) if t0 is not None else (lambda: 0.0)
    e = engine
    masses = np.array([0.0] + [systems.MASS[el] for el in systems.ELEMENTS])
    members = []    # [request, steps done, finished] in pack order
    live = False    # the context holds `members`
    closed = False
    cap_atoms = getattr(admit, "max_atoms", 4_000_000)
    while True:
        active = [m for m in members if not m[2]]
        fresh = []
        if not closed:
            got = admit(cap_atoms - sum(m[0].system.n for m in active), not active)
            if got is None:
                closed = True
            else:
                fresh = got
        if not active and not fresh:
            if closed:
                return
            continue
        waste = sum(m[0].system.n for m in members if m[2])
        total = sum(m[0].system.n for m in members) or 1
        if fresh or not live or waste > repack_waste * total:
            carried = []
            if live and active:   # the continuing members' state from the live context
                x, v, q = e.state()
                ms = e.md_state()
                offs = np.concatenate([[0], np.cumsum([m[0].system.n for m in members])])
                for k, m in enumerate(members):
                    if m[2]:
                        continue
                    a, b = offs[k], offs[k + 1]
                    carried.append((m, _Moved(m[0].system, x[a:b]), v[a:b], q[a:b], ms["s_hist"][:, a:b], ms["t_hist"][:, a:b],
                                    ms["nh"][k]))
            else:
                carried = [(m, m[0].system, None, None, None, None, None) for m in active]   # not started yet
            for r in fresh:
                r.start = now()
            new_members = [c[0] for c in carried] + [[r, 0, False] for r in fresh]
            sysl = [c[1] for c in carried] + [r.system for r in fresh]
            e.set_systems(sysl)
            e.qeq_setup(0.0, 10.0, e.extract("chi"), e.extract("eta"), e.extract("gamma"))
            reqs = [m[0] for m in new_members]
            v = np.concatenate([c[2] if c[2] is not None else systems.maxwell_boltzmann(c[0][0].system.masses(), c[0][0].t_target,
                                                                                           seed=c[0][0].seed) for c in carried] +
                               [systems.maxwell_boltzmann(r.system.masses(), r.t_target, seed=r.seed) for r in fresh])
            n_all = sum(sy.n for sy in sysl)
            q = np.zeros(n_all)
            sh, th = np.zeros((4, n_all)), np.zeros((4, n_all))
            nh = np.zeros((len(sysl), 32))
            keep = np.zeros(len(sysl), dtype=np.int32)
            a = 0
            for k, c in enumerate(carried):
                b = a + c[1].n
                if c[3] is not None:
                    keep[k] = 1
                    q[a:b] = c[3]
                    sh[:, a:b] = c[4]
                    th[:, a:b] = c[5]
                    nh[k] = c[6]
                a = b
            r0 = reqs[0]
            e.md_resume(v, masses, keep, q, {"s_hist": sh, "t_hist": th, "nh": nh}, dt=r0.dt, t_target=r0.t_target, t_damp=r0.t_damp,
                        skin=2.0, qeq_every=1, qeq_imax=200, qeq_tol=r0.qeq_tol)
            if any(r.t_target != r0.t_target or r.t_damp != r0.t_damp for r in reqs):
                e.md_system_targets([r.t_target for r in reqs], [r.t_damp for r in reqs])
            members = new_members
            live = True
        active = [m for m in members if not m[2]]
        k = min([chunk] + [m[0].steps - m[1] for m in active])
        thermo = e.md_run(k) if k > 0 else None
        for m in members:
            if not m[2]:
                m[1] += k
        done = [i for i, m in enumerate(members) if not m[2] and m[1] >= m[0].steps]
        if done:
            g = e.compute()["grad"]
            x, vel, q = e.state()
            per = e.system_results()
            offs = np.concatenate([[0], np.cumsum([m[0].system.n for m in members])])
            t = now()
            for i in done:
                r = members[i][0]
                r.result = _result(per, g, x, q, i, offs[i], offs[i + 1], thermo)
                r.end = t
                members[i][2] = True


def new_engine(ffield_text, precision):
    e = Engine(0)
    e.pair_style(["NULL", "precision", precision])
    e.pair_coeff(ffield_text, list(systems.ELEMENTS))
    return e


def engine_bytes_per_atom(e):
    """Device bytes per atom implied by an engine's current capacities (lists, slots, matrix, per-atom arrays)."""
    st = e.stats()
    if not st["n"]:
        return DEFAULT_BYTES_PER_ATOM
    per = 8 * st["maxnb"] + 6 * st["maxnb"] + 8 * st["maxnb"]      # skin list (j, entry), charge rows (col + value), FP64 copy
    per += 8 * 29 * st["maxb"] + 12 * st["maxb"]                     # bond slots
    per += 8 * 64 + 8 * 200                                          # per-atom arrays, candidate lists (typical)
    return per


class Server:
    """Concurrent executor of MD requests (see the module docstring). packed=False: each slot runs one request at a time in its
    own context. packed=True: a free slot takes the oldest queued request alone if it has at least pack_below atoms (the
    single-system path, with cell lists), else every queued request below that size that fits the memory budget (at most
    max_pack_atoms atoms), packed in one context (run_pack_chunks). continuous=True: small requests are packed only when
    capacity is tight: while another slot is idle, a lone small request runs alone on the single-system path; otherwise a free
    slot starts a continuous pack (run_continuous) of the queued small requests, which, while every slot is busy, also admits
    the small requests that arrive, at its chunk boundaries (every `chunk` steps), instead of leaving them queued until a slot
    frees (with an idle slot, the arrival takes that slot instead, which costs no re-pack)."""

    def __init__(self, ffield_text, slots=4, precision="mixed", mem_budget_bytes=200 << 30, packed=False, max_pack_atoms=4_000_000,
                 pack_below=3000, continuous=False, chunk=50):
        self.ffield_text = ffield_text
        self.precision = precision
        self.slots = slots
        self.packed = packed or continuous
        self.continuous = continuous   # packed mode with admission into running packs at chunk boundaries (run_continuous)
        self.chunk = chunk
        self.max_pack_atoms = max_pack_atoms
        self.pack_below = pack_below   # packed mode: requests of at least this many atoms run alone on the single-system path
        self.budget = mem_budget_bytes
        self.bytes_per_atom = DEFAULT_BYTES_PER_ATOM
        self._free = queue.Queue()
        self._lock = threading.Condition()
        self._mem_in_use = 0
        self._engines = [new_engine(ffield_text, precision) for _ in range(slots)]
        for k in range(slots):
            self._free.put(k)

    def close(self):
        for e in self._engines:
            e.close()

    def _need(self, req):
        return req.system.n * self.bytes_per_atom

    def run(self, requests, realtime=False):
        """Runs `requests` (in arrival order). realtime=True honours Request.arrival (seconds after start); otherwise all
        requests are queued at once. Returns the requests with start/end times (seconds after start) and results."""
        if self.packed:
            return self._run_packed(requests, realtime)
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        threads = []
        errors = []
        admitted = {}   # device bytes reserved for each running request

        def work(slot, req):
            try:
                req.result = run_one(self._engines[slot], self.ffield_text, req, self.precision)
                self.bytes_per_atom = max(self.bytes_per_atom, engine_bytes_per_atom(self._engines[slot]))
            except Exception as ex:   # surfaced after the run
                errors.append((req.rid, ex))
            finally:
                req.end = time.perf_counter() - t0
                with self._lock:
                    self._mem_in_use -= admitted[req.rid]
                    self._lock.notify_all()
                self._free.put(slot)

        for req in reqs:
            if realtime:
                wait = req.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            need = self._need(req)
            if need > self.budget:
                raise ValueError("request %d needs more device memory than the budget" % req.rid)
            with self._lock:   # admission by device memory
                while self._mem_in_use + need > self.budget:
                    self._lock.wait()
                self._mem_in_use += need
                admitted[req.rid] = need
            slot = self._free.get()   # admission by slot
            req.start = time.perf_counter() - t0
            th = threading.Thread(target=work, args=(slot, req), daemon=True)
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        if errors:
            rid, ex = errors[0]
            raise RuntimeError("request %d failed: %s" % (rid, ex))
        return reqs


    def _run_packed(self, requests, realtime):
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        waiting = []
        state = {"submitted": 0, "idle": 0}
        cv = threading.Condition()
        errors = []

        def take():   # the oldest request alone if it is large, else every queued small request that fits (arrival order)
            first = waiting[0]
            if first.system.n * self.bytes_per_atom > self.budget:
                raise ValueError("request %d needs more device memory than the budget" % first.rid)
            if first.system.n >= self.pack_below:
                return [waiting.pop(0)]
            pack, atoms, k = [], 0, 0
            while k < len(waiting):
                r = waiting[k]
                if r.system.n >= self.pack_below:
                    k += 1
                    continue
                if pack and (atoms + r.system.n > self.max_pack_atoms or (atoms + r.system.n) * self.bytes_per_atom > self.budget):
                    break
                pack.append(waiting.pop(k))
                atoms += r.system.n
            return pack

        def admit_small(capacity, block):   # continuous packs: the queued small requests that fit (arrival order)
            with cv:
                while block and not waiting and state["submitted"] < len(reqs):
                    cv.wait()
                if not block and state["idle"] > 0:
                    return []   # an idle slot starts its own pack: a running one re-packs only when every slot is busy
                out, atoms, k = [], 0, 0
                while k < len(waiting):
                    r = waiting[k]
                    if r.system.n >= self.pack_below or (out and atoms + r.system.n > capacity):
                        k += 1
                        continue
                    if r.system.n * self.bytes_per_atom > self.budget:
                        raise ValueError("request %d needs more device memory than the budget" % r.rid)
                    out.append(waiting.pop(k))
                    atoms += r.system.n
                if block and not out:
                    return None   # nothing small queued: the worker returns to the queue
                return out
        admit_small.max_atoms = self.max_pack_atoms

        def worker(slot):
            while True:
                with cv:
                    state["idle"] += 1
                    while not waiting and state["submitted"] < len(reqs):
                        cv.wait()
                    state["idle"] -= 1
                    if not waiting:
                        return
                    try:
                        if self.continuous and waiting[0].system.n < self.pack_below:
                            # while other slots are idle a lone small request runs alone on the single-system path; with
                            # capacity tight (no other idle slot, or several small requests queued) a continuous pack admits
                            # the queued small requests itself, and later arrivals join it at chunk boundaries
                            nsmall = sum(1 for r in waiting if r.system.n < self.pack_below)
                            pack = [waiting.pop(0)] if nsmall == 1 and state["idle"] > 0 else None
                        else:
                            pack = take()
                    except Exception as ex:
                        errors.append((waiting[0].rid, ex))
                        waiting.clear()
                        return
                try:
                    if pack is None:
                        run_continuous(self._engines[slot], self.ffield_text, admit_small, chunk=self.chunk, t0=t0)
                    elif len(pack) == 1 and (pack[0].system.n >= self.pack_below or self.continuous):
                        r = pack[0]
                        r.start = time.perf_counter() - t0
                        r.result = run_one(self._engines[slot], self.ffield_text, r, self.precision)
                        r.end = time.perf_counter() - t0
                    else:
                        run_pack_chunks(self._engines[slot], self.ffield_text, pack, t0)
                    self.bytes_per_atom = max(self.bytes_per_atom, engine_bytes_per_atom(self._engines[slot]))
                except Exception as ex:   # surfaced after the run
                    errors.append((pack[0].rid if pack else -1, ex))

        threads = [threading.Thread(target=worker, args=(k,), daemon=True) for k in range(self.slots)]
        for th in threads:
            th.start()
        for r in reqs:
            if realtime:
                wait = r.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            with cv:
                waiting.append(r)
                state["submitted"] += 1
                cv.notify_all()
        with cv:
            cv.notify_all()
        for th in threads:
            th.join()
        if errors:
            rid, ex = errors[0]
            raise RuntimeError("request %d failed: %s" % (rid, ex))
        return reqs


def batch_systems(count=512, nmin=100, nmax=3000, seed=1):
    """The batched workload: `count` periodic systems, N log-uniform in [nmin, nmax]; even indices water (216-molecule recipe),
    odd indices CH4:O2 = 1:2 gas at 0.3 g/cm^3 with 2.5 A between molecules."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        if k % 2 == 0:
            side = max(2, int(round((n / 3.0) ** (1.0 / 3.0))))
            out.append(systems.water_box(side, seed=seed * 100003 + k))
        else:
            out.append(systems.ch4_o2_gas(int(n), density=0.3, seed=seed * 100003 + k))
    return out


def latency_stats(reqs):
    lat = np.array([r.end - r.arrival for r in reqs])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean())}


def serving_requests(count=32, nmin=200, nmax=20000, smin=10, smax=1000, seed=2, t_target=2500.0, dt=0.1):
    """The serving workload: systems from the batch generator with N log-uniform in [nmin, nmax] and step counts uniform in
    [smin, smax]; arrival times are set by the caller (e.g. Poisson)."""
    rng = np.random.default_rng(seed)
    syst = batch_systems(count, nmin=nmin, nmax=nmax, seed=seed)
    return [Request(rid=k, system=sy, steps=int(rng.integers(smin, smax + 1)), t_target=t_target, dt=dt, seed=1000 + k)
            for k, sy in enumerate(syst)]
