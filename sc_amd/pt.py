"""Batched parallel tempering on one GPU.

The reference runs one replica per MPI process and exchanges temperatures
between neighbours of its ladder every `nrepchange` sweeps
(MoveCreator::replicaExchangeMove, movecreator.cpp:555-760; alternating
even/odd pairs, :607-617).  Here every replica lives in one context: all of
them are swept together, one kernel launch per checkerboard colour pass, and
the exchange step runs on the whole-system energies of all replicas.

`c3_batch` builds the synthetic batched-tempering system: 64 replicas of
uneven size (N uniform in 500-8000), each a 50/50 PSC + CPSC mixture (L = 3)
at a packing fraction uniform in 0.15-0.35, temperatures geometric over
0.8-1.6 T*.
"""
import numpy as np

from . import workloads as W


def c3_spec(nrep=64, nmin=500, nmax=8000, eta=(0.15, 0.35), tmin=0.8, tmax=1.6, seed=3):
    """Per replica: particle count, packing fraction, temperature, box edge."""
    rng = np.random.default_rng(seed)
    n = rng.integers(nmin, nmax + 1, nrep)
    n += n % 2   # even: half PSC, half CPSC
    et = rng.uniform(eta[0], eta[1], nrep)
    temps = tmin * (tmax / tmin) ** (np.arange(nrep) / max(1, nrep - 1))
    v = W.particle_volume(1.0, 3.0)
    box = (n * v / et) ** (1.0 / 3.0)
    return n.astype(np.int32), et, temps, box


def c3_topology(total):
    half = total // 2
    return W.top_init([W.TYPES["psc"], W.TYPES["cpsc"]], [half, total - half])


def c3_batch(g, nrep=64, seed=3, eta_rsa=0.2, steps=8, sweeps_per_step=3, relax=10, transmx=0.3,
             rot_half=20.0 / 180 * np.pi / 4, nmin=500, nmax=8000, log=None):
    """Load the batch into GpuSystem g: random sequential addition of every
    replica at min(eta, eta_rsa), then the replicas denser than that are
    compressed in `steps` isotropic steps with sweeps between them.  Returns
    (n per replica, temperatures, boxes)."""
    n, et, temps, box = c3_spec(nrep, nmin, nmax, seed=seed)
    total = int(n.sum())
    g.load_topology(c3_topology(total))
    box0 = box * (et / np.minimum(et, eta_rsa)) ** (1.0 / 3.0)
    pos, d, q, types = [], [], [], []
    for k in range(nrep):
        t = np.array([1] * (n[k] // 2) + [2] * (n[k] - n[k] // 2), np.int32)
        p_, d_, q_ = g.generate_rsa(t, np.full(3, box0[k]), seed + 17 * k)
        pos.append(p_)
        d.append(d_)
        q.append(q_)
        types.append(t)
    g.set_config(np.full(3, box0[0]), np.concatenate(pos), np.concatenate(d), np.concatenate(q),
                 np.concatenate(types))
    boxes = np.repeat(box0[:, None], 3, 1)
    g.batch_set(n, boxes)
    counter = 10 ** 9
    for s in range(1, steps + 1):
        b = box0 * (box / box0) ** (s / steps)
        boxes = np.repeat(b[:, None], 3, 1)
        g.batch_set(n, boxes)
        g.batch_sweep(sweeps_per_step, temps, transmx, rot_half, r=2, seed=seed, counter=counter)
        counter += sweeps_per_step
        if log:
            log("c3 compress %d/%d" % (s, steps))
    g.batch_sweep(relax, temps, transmx, rot_half, r=2, seed=seed, counter=counter)
    return n, temps, np.repeat(box[:, None], 3, 1)


def run(g, temps, nsweeps, exchange_every=10, transmx=0.3, rot_half=20.0 / 180 * np.pi / 4, seed=1, counter=0,
        r=0, refresh_every=10):
    """nsweeps sweeps of the batch in g with a replica-exchange step every
    `exchange_every` sweeps (even/odd ladder pairs alternating as the
    reference does).  The exchange step reads each replica's whole-system
    energy carried through the sweeps (the FP64 sum of its accepted changes:
    temperatures move between replicas, configurations stay), evaluated in
    full at the first exchange and at every `refresh_every`-th one after it
    (1: at every exchange).  Returns (temperatures, sweep stats summed per
    replica, accepted and attempted exchanges per ladder pair)."""
    temps = np.array(temps, np.float64)
    nrep = len(temps)
    acc = np.zeros(max(nrep - 1, 1), np.int64)
    tried = np.zeros(max(nrep - 1, 1), np.int64)
    tot = [dict() for _ in range(nrep)]
    done = 0
    block = 0
    e = None
    while done < nsweeps:
        k = min(exchange_every, nsweeps - done)
        st = g.batch_sweep(k, temps, transmx, rot_half, r=r, seed=seed, counter=counter + done)
        for t, s in zip(tot, st):
            for key, v in s.items():
                t[key] = t.get(key, 0) + v
        if e is not None:
            e = e + np.array([s["energy_change"] for s in st], np.float64)
        done += k
        if k == exchange_every:
            if e is None or block % max(1, int(refresh_every)) == 0:
                e = g.batch_energies()
            parity = block % 2
            temps, a = g.batch_exchange(temps, e, parity, seed=seed, counter=counter + done)
            acc[: nrep - 1] += a
            tried[parity: nrep - 1: 2] += 1
            block += 1
    return temps, tot, acc[: nrep - 1], tried[: nrep - 1]
