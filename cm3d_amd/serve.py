"""Continuous-batching MD server on one GPU.

cm3d runs one system per process (main/md_main.new.c:61-134); here many
independent MD jobs -- each a system with its own size, box, cutoffs and step
count -- share the device.  A job is admitted when it has arrived, a slot is
free and its device memory fits the budget; it then keeps coordinates,
velocities, lists and tables resident (an NBContext + NBMD) until its last
step, when its final state is copied out and its memory released.  All
admitted jobs advance together: every kernel of the step is launched once for
the whole set (NBBatch, replayed from a hipGraph), so jobs that arrive while
others run join the next batch and their work overlaps on the device.

A job's trajectory in the mix is bitwise the one it follows alone: every
system's arithmetic and summation order are the same in and out of a batch,
and each job runs exactly its own step count (a batch advances by the fewest
steps any member has left, capped at tick_steps).

Timing: per-step latency is the wall time of a batched step (a tick of k
steps, launched and waited for, divided by k); per-job latency is completion
minus arrival (queueing included).
"""
import ctypes as C
import math
import time
from collections import deque

import numpy as np

from . import nb

_HIP = None


def _device_sync():
    global _HIP
    if _HIP is None:
        nb.lib()  # the HIP runtime is loaded with the library
        _HIP = C.CDLL("libamdhip64.so")
    rc = _HIP.hipDeviceSynchronize()
    if rc != 0:
        raise nb.NBError(f"hipDeviceSynchronize failed ({rc})")


def _ceil(a, b):
    return -(-a // b)


def estimate_bytes(sysd, sharing=1):
    """Device bytes a job will need, before its context exists, following
    the allocation sizes of nb_create / nb_md_create: the neighbour lists
    (capacity 1.6x the mean full-list length + 64 entries, grown by 1.25x on
    the rare overflow -- allowed for here), per-atom state, and the Ewald
    buffers (rows of consecutive kc, structure-factor lanes of <= 16 values
    of |kc|, structure-factor chunks and force ranges sized as nb_create
    sizes them)."""
    n = int(sysd["natoms"])
    h = np.asarray(sysd["hmat"], float).reshape(3, 3)
    vol = abs(np.linalg.det(h))
    r2 = float(np.max(sysd["rmaxl_skin2"])) if len(sysd["rmaxl_skin2"]) else 0.
    nq = int(np.count_nonzero(sysd["qch"]))
    if nq:
        r2 = max(r2, float(np.asarray(sysd["rmaxl_skin2_e"]).reshape(-1)[0]))
    cap = int(1.25 * (1.6 * 4.18879 * r2 ** 1.5 * n / vol + 64)) + 128
    ktot = int(np.asarray(sysd.get("ktot", 0)).reshape(-1)[0]) if nq else 0
    nrow = nsym = nent = 0
    if ktot:
        ka, kb, kc = (np.asarray(sysd[k]) for k in ("ka", "kb", "kc"))
        lo = 0
        for k in range(1, ktot + 1):  # rows: runs of consecutive kc at one (ka, kb)
            if k < ktot and ka[k] == ka[k - 1] and kb[k] == kb[k - 1] and kc[k] == kc[k - 1] + 1:
                continue
            a, b = int(kc[lo]), int(kc[k - 1])
            mlo = a if a > 0 else (-b if b < 0 else 0)
            cnt = max(abs(a), abs(b)) - mlo + 1  # values of m = |kc|
            nrow += 1
            nent += cnt
            nsym += _ceil(cnt, 16)
            lo = k
    nsw = max(1, _ceil(nsym, 64))
    share = max(1, int(sharing))
    nch = _ceil(max(1, 4096 // share), nsw)
    nrange = _ceil(max(1, 2048 // share), max(1, _ceil(nq, 768)))
    if share > 1:  # a bounded k x charge product per workgroup (65536 each pass)
        nch = max(nch, _ceil(ktot * nq, 65536 * nsw))
        nrange = max(nrange, _ceil(ktot * 768, 65536))
    nch = max(1, min(nch, _ceil(nq, 64)))
    nrange = max(1, min(nrange, max(1, min(nrow, nent // 64))))
    per_atom = 4 * cap + 516           # lists + positions, forces, records, MD state
    #                                    (with the batch tick's 96 B snapshot)
    ewald = nq * (64 + 24 * nrange) + ktot * (16 * nch + 32) + nrow * 36 + nent * 16 + nsym * 32
    ncells = (int(2 * (vol ** (1. / 3.)) / math.sqrt(max(r2, 1e-12))) + 1) ** 3
    return n * per_atom + ewald + 16 * ncells + (1 << 20)


class Job:
    """One MD job: a system dict (cm3d_amd.systems) -- or a zero-argument
    `make` that builds it, called when the job is prepared -- its initial
    velocities and step count, its arrival time (s from the start of the
    run)."""

    def __init__(self, jid, sysd, nsteps, arrival, nres_intra=1, dt_ps=0.002, vel=None,
                 sharing=16, make=None, vel_seed=None, temp=300.0):
        self.id = jid
        self.sys = sysd
        self.make = make
        self.nsteps = int(nsteps)
        self.arrival = float(arrival)
        self.nres_intra = int(nres_intra)
        self.dt_ps = float(dt_ps)
        self.vel = vel
        self.vel_seed = vel_seed
        self.temp = float(temp)
        self.sharing = int(sharing)
        self.natoms = int(sysd["natoms"]) if sysd is not None else 0
        self.est_bytes = None
        self.ctx = self.md = None
        self.done = 0
        self.bytes = 0
        self.t_admit = self.t_done = None
        self.result = None
        self.failed = None
        if sysd is not None:
            self.prepare()

    def prepare(self):
        """Host-side set-up (system, velocities, footprint estimate): no GPU
        work, so a server does it ahead of admission, off its tick loop."""
        if self.sys is None:
            self.sys = self.make()
            self.natoms = int(self.sys["natoms"])
        if self.vel is None:
            if self.vel_seed is None:
                self.vel = self.sys["vel"]
            else:
                from . import systems as S
                self.vel = S.maxwell_velocities(self.sys["amass"], self.temp, self.vel_seed)
        if self.est_bytes is None:
            self.est_bytes = estimate_bytes(self.sys, self.sharing)
        return self

    def start(self, precision):
        self.ctx = nb.NBContext(self.sys, precision=precision, sharing=self.sharing)
        self.md = nb.NBMD(self.ctx, self.sys, self.dt_ps, nres_intra=self.nres_intra)
        self.md.set_state(self.sys["pos"], self.vel, self.sys["hmat"])
        self.md.init()
        self.bytes = self.ctx.footprint()

    def finish(self, keep_result=True):
        pos, vel = self.md.get_state()
        # cm3d stops a run whose energies turn non-finite (exit 99,
        # output/md_writescreen.c:82-90); here only this job fails
        if not (np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))):
            self.failed = "non-finite positions or velocities"
        self.result = (pos, vel) if keep_result else None
        self.md.close()
        self.ctx.close()
        self.md = self.ctx = None
        if self.make is not None:
            self.sys = None  # rebuilt on demand; the result is what is kept


class Server:
    """Admission by arrival, free slot and HBM budget (FIFO); one batch of
    all admitted jobs, re-recorded when membership changes.

    A tick (a batched step of k steps) runs on a worker thread while the
    loop admits the jobs that arrived meanwhile (their contexts are set up
    on their own streams, beside the running tick) and a second worker
    prepares the next jobs' systems on the host; admitted jobs join at the
    next tick.  A tick whose list builds overflowed redoes the systems
    concerned with more slots (nb_batch_step), and a job whose state turns
    non-finite fails alone."""

    def __init__(self, budget_bytes, max_slots=64, tick_steps=50, precision="mixed",
                 graph=True, keep_results=True, lookahead=32):
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        self.tick_steps = int(tick_steps)
        self.precision = precision
        self.graph = graph
        self.keep_results = keep_results
        self.lookahead = int(lookahead)

    def run(self, jobs, log=None):
        from concurrent.futures import ThreadPoolExecutor
        order = sorted(jobs, key=lambda j: (j.arrival, j.id))
        pending = deque(order)
        prep_pool = ThreadPoolExecutor(1)
        tick_pool = ThreadPoolExecutor(1)
        prepared = {}
        nprep = 0

        def top_up():
            nonlocal nprep
            while nprep < len(order) and nprep < len(order) - len(pending) + self.lookahead:
                j = order[nprep]
                prepared[j.id] = prep_pool.submit(j.prepare)
                nprep += 1

        active, joining, used = [], [], 0
        batch, dirty = None, False
        step_lat, step_w, batch_sizes, rebuilds = [], [], [], 0
        peak_bytes, peak_slots, recoveries = 0, 0, 0
        t0 = time.perf_counter()

        def admit():
            nonlocal used
            top_up()
            n = 0
            while (pending and pending[0].arrival <= time.perf_counter() - t0
                   and len(active) + len(joining) < self.max_slots):
                j = pending[0]
                prepared[j.id].result()  # host set-up done (raises its error)
                if (active or joining) and used + j.est_bytes > self.budget:
                    break  # head of the queue waits for memory
                pending.popleft()
                del prepared[j.id]
                j.start(self.precision)
                j.t_admit = time.perf_counter() - t0
                used += j.bytes
                joining.append(j)
                n += 1
            top_up()
            return n

        try:
            while pending or active or joining:
                admit()
                if joining:
                    active += joining
                    joining = []
                    dirty = True
                peak_bytes = max(peak_bytes, used)
                peak_slots = max(peak_slots, len(active))
                if not active:
                    if pending:
                        time.sleep(max(0., min(0.01, pending[0].arrival -
                                               (time.perf_counter() - t0))))
                    continue
                if dirty:
                    if batch is not None:
                        recoveries += batch.recoveries()
                        batch.close()
                    batch = nb.NBBatch([j.md for j in active])
                    rebuilds += 1
                    dirty = False
                k = min(self.tick_steps, min(j.nsteps - j.done for j in active))
                t = time.perf_counter()
                # re-recording and graph capture happen here, between ticks:
                # a capture forbids the legacy-stream copies of the
                # admissions that overlap the tick on this thread
                batch.step(0, None, self.graph)
                fut = tick_pool.submit(batch.step, k, None, self.graph)
                while not fut.done():
                    if not admit():  # admissions overlap the running tick
                        time.sleep(2e-4)
                fut.result()
                dt = time.perf_counter() - t
                step_lat.append(dt / k)
                step_w.append(k)
                batch_sizes.append(len(active))
                for j in active:
                    j.done += k
                fin = [j for j in active if j.done >= j.nsteps]
                if fin:
                    recoveries += batch.recoveries()
                    batch.close()
                    batch = None
                    for j in fin:
                        j.finish(self.keep_results)
                        j.t_done = time.perf_counter() - t0
                        used -= j.bytes
                    active = [j for j in active if j.done < j.nsteps]
                    dirty = True
                    if log:
                        log(f"t={j.t_done:.2f}s done {len(fin)} job(s); active {len(active)}, "
                            f"pending {len(pending)}")
        finally:
            tick_pool.shutdown(wait=True)
            prep_pool.shutdown(wait=True)
        wall = time.perf_counter() - t0
        if batch is not None:
            recoveries += batch.recoveries()
            batch.close()
        lat = np.repeat(np.array(step_lat), np.array(step_w)) if step_lat else np.zeros(1)
        jl = np.array([j.t_done - j.arrival for j in jobs])
        atom_steps = float(sum(j.natoms * j.nsteps for j in jobs))
        return dict(wall_s=wall, atom_steps=atom_steps, atom_steps_per_s=atom_steps / wall,
                    jobs=len(jobs), ticks=len(step_lat), batch_rebuilds=rebuilds,
                    mean_batch=float(np.average(batch_sizes, weights=step_w)) if step_w else 0.,
                    peak_slots=peak_slots, peak_bytes=peak_bytes,
                    overflow_recoveries=int(recoveries),
                    failed=[j.id for j in jobs if j.failed],
                    p50_step_ms=1e3 * float(np.percentile(lat, 50)),
                    p99_step_ms=1e3 * float(np.percentile(lat, 99)),
                    p50_job_s=float(np.percentile(jl, 50)), p99_job_s=float(np.percentile(jl, 99)))


def job_stream(njobs, rate_per_s, seed, water_nside=(7, 20), mix_atoms=(500, 20000),
               steps=(200, 5000)):
    """Poisson arrivals (mean rate_per_s) of jobs drawn half from the LJ +
    charge binary-mixture family (N log-uniform in mix_atoms) and half from
    flexible water (nside^3 molecules: 343-8000, the cubes within 256-8192),
    step counts uniform in steps.  Every job is its own random instance: its
    own placement seed, size and Maxwell-Boltzmann velocities at 300 K; the
    system is built when the job is prepared."""
    from . import systems as S
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(1.0 / rate_per_s, njobs))
    jobs = []
    for i in range(njobs):
        sseed = int(rng.integers(1 << 30))
        if rng.random() < 0.5:
            n = int(math.exp(rng.uniform(math.log(mix_atoms[0]), math.log(mix_atoms[1]))))
            make = (lambda n=n, sd=sseed: S.mixture(n, seed=sd, rc=9.0, kmax=7))
            nri = 1
        else:
            side = int(rng.integers(water_nside[0], water_nside[1] + 1))
            make = (lambda side=side, sd=sseed: S.water(nside=side, seed=sd))
            nri = 4
        nst = int(rng.integers(steps[0], steps[1] + 1))
        jobs.append(Job(i, None, nst, float(arr[i]), nres_intra=nri, make=make,
                        vel_seed=int(rng.integers(1 << 30))))
    return jobs
