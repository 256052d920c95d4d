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
    buffers (segments of <= 16 consecutive kc, structure-factor chunks and
    force ranges sized as nb_create sizes them)."""
    n = int(sysd["natoms"])
    h = np.asarray(sysd["hmat"], float).reshape(3, 3)
    vol = abs(np.linalg.det(h))
    r2 = float(np.max(sysd["rmaxl_skin2"])) if len(sysd["rmaxl_skin2"]) else 0.
    nq = int(np.count_nonzero(sysd["qch"]))
    if nq:
        r2 = max(r2, float(np.asarray(sysd["rmaxl_skin2_e"]).reshape(-1)[0]))
    cap = int(1.25 * (1.6 * 4.18879 * r2 ** 1.5 * n / vol + 64)) + 128
    ktot = int(np.asarray(sysd.get("ktot", 0)).reshape(-1)[0]) if nq else 0
    nseg = 0
    if ktot:
        ka, kb, kc = (np.asarray(sysd[k]) for k in ("ka", "kb", "kc"))
        run = 0
        for k in range(ktot):
            cont = k > 0 and ka[k] == ka[k - 1] and kb[k] == kb[k - 1] and kc[k] == kc[k - 1] + 1
            if cont and run < 16:
                run += 1
            else:
                nseg += 1
                run = 1
    nsw = max(1, _ceil(nseg, 64))
    share = max(1, int(sharing))
    nch = _ceil(max(1, 4096 // share), nsw)
    nrange = _ceil(max(1, 1024 // share), max(1, _ceil(nq, 512)))
    if share > 1:  # a bounded k x charge product per workgroup (65536 each pass)
        nch = max(nch, _ceil(ktot * nq, 65536 * nsw))
        nrange = max(nrange, _ceil(ktot * 512, 65536))
    nch = max(1, min(nch, _ceil(nq, 32)))
    nrange = max(1, min(nrange, max(1, nseg)))
    per_atom = 4 * cap + 420           # lists + positions, forces, records, MD state
    ewald = nq * (64 + 24 * nrange) + ktot * (16 * nch + 96) + nseg * 48
    ncells = (int(2 * (vol ** (1. / 3.)) / math.sqrt(max(r2, 1e-12))) + 1) ** 3
    return n * per_atom + ewald + 16 * ncells + (1 << 20)


class Job:
    """One MD job: a system dict (cm3d_amd.systems), its initial velocities
    and step count, its arrival time (s from the start of the run)."""

    def __init__(self, jid, sysd, nsteps, arrival, nres_intra=1, dt_ps=0.002, vel=None,
                 sharing=16):
        self.id = jid
        self.sys = sysd
        self.nsteps = int(nsteps)
        self.arrival = float(arrival)
        self.nres_intra = int(nres_intra)
        self.dt_ps = float(dt_ps)
        self.vel = sysd["vel"] if vel is None else vel
        self.natoms = int(sysd["natoms"])
        self.sharing = int(sharing)
        self.est_bytes = estimate_bytes(sysd, self.sharing)
        self.ctx = self.md = None
        self.done = 0
        self.bytes = 0
        self.t_admit = self.t_done = None
        self.result = None

    def start(self, precision):
        self.ctx = nb.NBContext(self.sys, precision=precision, sharing=self.sharing)
        self.md = nb.NBMD(self.ctx, self.sys, self.dt_ps, nres_intra=self.nres_intra)
        self.md.set_state(self.sys["pos"], self.vel, self.sys["hmat"])
        self.md.init()
        self.bytes = self.ctx.footprint()

    def finish(self):
        self.result = self.md.get_state()
        self.md.close()
        self.ctx.close()
        self.md = self.ctx = None


class Server:
    """Admission by arrival, free slot and HBM budget (FIFO); one batch of
    all admitted jobs, rebuilt when membership changes."""

    def __init__(self, budget_bytes, max_slots=64, tick_steps=50, precision="mixed",
                 graph=True):
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        self.tick_steps = int(tick_steps)
        self.precision = precision
        self.graph = graph

    def run(self, jobs, log=None):
        pending = deque(sorted(jobs, key=lambda j: (j.arrival, j.id)))
        active, used = [], 0
        batch, dirty = None, False
        step_lat, step_w, batch_sizes, rebuilds = [], [], [], 0
        peak_bytes, peak_slots = 0, 0
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now and len(active) < self.max_slots:
                j = pending[0]
                if active and used + j.est_bytes > self.budget:
                    break  # head of the queue waits for memory
                pending.popleft()
                j.start(self.precision)
                j.t_admit = time.perf_counter() - t0
                used += j.bytes
                active.append(j)
                dirty = True
            peak_bytes = max(peak_bytes, used)
            peak_slots = max(peak_slots, len(active))
            if not active:
                if pending:
                    time.sleep(max(0., pending[0].arrival - (time.perf_counter() - t0)))
                continue
            if dirty:
                if batch is not None:
                    batch.close()
                batch = nb.NBBatch([j.md for j in active])
                rebuilds += 1
                dirty = False
            k = min(self.tick_steps, min(j.nsteps - j.done for j in active))
            t = time.perf_counter()
            batch.step(k, graph=self.graph)
            _device_sync()
            dt = time.perf_counter() - t
            step_lat.append(dt / k)
            step_w.append(k)
            batch_sizes.append(len(active))
            for j in active:
                j.done += k
            fin = [j for j in active if j.done >= j.nsteps]
            if fin:
                batch.close()
                batch = None
                for j in fin:
                    j.finish()
                    j.t_done = time.perf_counter() - t0
                    used -= j.bytes
                active = [j for j in active if j.done < j.nsteps]
                dirty = True
                if log:
                    log(f"t={j.t_done:.2f}s done {len(fin)} job(s); active {len(active)}, "
                        f"pending {len(pending)}")
        wall = time.perf_counter() - t0
        if batch is not None:
            batch.close()
        lat = np.repeat(np.array(step_lat), np.array(step_w)) if step_lat else np.zeros(1)
        jl = np.array([j.t_done - j.arrival for j in jobs])
        atom_steps = float(sum(j.natoms * j.nsteps for j in jobs))
        return dict(wall_s=wall, atom_steps=atom_steps, atom_steps_per_s=atom_steps / wall,
                    jobs=len(jobs), ticks=len(step_lat), batch_rebuilds=rebuilds,
                    mean_batch=float(np.average(batch_sizes, weights=step_w)) if step_w else 0.,
                    peak_slots=peak_slots, peak_bytes=peak_bytes,
                    p50_step_ms=1e3 * float(np.percentile(lat, 50)),
                    p99_step_ms=1e3 * float(np.percentile(lat, 99)),
                    p50_job_s=float(np.percentile(jl, 50)), p99_job_s=float(np.percentile(jl, 99)))


def job_stream(njobs, rate_per_s, seed, water_nside=(7, 20), mix_atoms=(500, 20000),
               steps=(200, 5000), ntemplates=24):
    """Poisson arrivals (mean rate_per_s) of jobs drawn half from the LJ +
    charge binary-mixture family (N log-uniform in mix_atoms) and half from
    flexible water (nside^3 molecules: 343-8000, the cubes within 256-8192),
    step counts uniform in steps.  Systems
    come from a pool of ntemplates generated up front; every job gets its own
    Maxwell-Boltzmann velocities at 300 K."""
    from . import systems as S
    rng = np.random.default_rng(seed)
    temps = []
    for t in range(ntemplates):
        if t % 2 == 0:
            n = int(math.exp(rng.uniform(math.log(mix_atoms[0]), math.log(mix_atoms[1]))))
            temps.append((S.mixture(n, seed=1000 + t, rc=9.0, kmax=7), 1))
        else:
            side = int(rng.integers(water_nside[0], water_nside[1] + 1))
            temps.append((S.water(nside=side, seed=2000 + t), 4))
    arr = np.cumsum(rng.exponential(1.0 / rate_per_s, njobs))
    jobs = []
    for i in range(njobs):
        sysd, nri = temps[int(rng.integers(ntemplates))]
        vel = S.maxwell_velocities(sysd["amass"], 300.0, int(rng.integers(1 << 30)))
        nst = int(rng.integers(steps[0], steps[1] + 1))
        jobs.append(Job(i, sysd, nst, float(arr[i]), nres_intra=nri, vel=vel))
    return jobs
