"""The learned SCHEFF field surrogate: the network wrapper (hps_nn_*), its
normalisation, training-data capture from the engine's own SCHEFF solve in
batched tracking, and training.

The grouped field solve of a batch (Batch.set_field_mode('gemm')) puts every
beam's mesh on an aspect grid, so its field is

    E = (c1 / dr^2) * G(kappa) n

with n the particles per bin, kappa = dz gamma / dr the cell aspect and G the
Green's operator normalised to dr = 1 (space_charge_kernel.cu:38-100).  The
network learns G(kappa) n / Q (Q = sum of n) from x = n nr nz / Q, with the
conditioning log(kappa); its output is divided by out_scale (the RMS of the
training targets) and log(kappa) is standardised by (u_mu, u_sigma).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import check

DEFAULT_WIDTHS = (48, 64, 96, 96)
DEFAULT_NBOT = 4


def _lib():
    return N.lib()


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Network:
    """The U-Net surrogate for an nr x nz SCHEFF mesh."""

    def __init__(self, nr=32, nz=128, widths=DEFAULT_WIDTHS, nbot=DEFAULT_NBOT, max_batch=64,
                 seed=1):
        h = C.c_void_p()
        w = (C.c_uint32 * 4)(*[int(v) for v in widths])
        check(_lib().hps_nn_create(C.byref(h), int(nr), int(nz), w, int(nbot), int(max_batch),
                                   int(seed)))
        self._h = h
        self.max_batch = int(max_batch)
        d = (C.c_uint32 * 7)()
        check(_lib().hps_nn_dims(h, d))
        self.Z0, self.cin, self.cout, self.cpad, self.zn, self.nr, self.nz = (int(v) for v in d)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._lib is not None:
            N._lib.hps_nn_destroy(h)
            self._h = None

    @property
    def num_params(self):
        return _lib().hps_nn_num_params(self._h)

    def params(self):
        p = np.empty(self.num_params, np.float32)
        check(_lib().hps_nn_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.num_params
        check(_lib().hps_nn_set_params(self._h, _fp(p)))

    def layers(self):
        n = C.c_uint32()
        check(_lib().hps_nn_layers(self._h, None, 0, C.byref(n)))
        t = (C.c_int32 * (8 * n.value))()
        check(_lib().hps_nn_layers(self._h, t, n.value, C.byref(n)))
        return [tuple(t[8 * k:8 * k + 8]) for k in range(n.value)]

    def set_norm(self, out_scale, u_mu, u_sigma):
        check(_lib().hps_nn_set_norm(self._h, float(out_scale), float(u_mu), float(u_sigma)))

    def norm(self):
        v = np.zeros(3, np.float32)
        check(_lib().hps_nn_get_norm(self._h, _fp(v)))
        return tuple(float(a) for a in v)

    # dataset ------------------------------------------------------------------
    def reserve(self, capacity):
        check(_lib().hps_nn_data_reserve(self._h, int(capacity)))

    def size(self):
        n = C.c_uint32()
        check(_lib().hps_nn_data_size(self._h, C.byref(n)))
        return n.value

    def shuffle(self, seed=0):
        """Random order of the dataset samples."""
        check(_lib().hps_nn_data_shuffle(self._h, int(seed)))

    def append(self, x, y, log_aspect):
        x, y, u = _f32(x), _f32(y), _f32(log_aspect)
        check(_lib().hps_nn_data_append(self._h, _fp(x), _fp(y), _fp(u), len(u)))

    def get(self, first, count, x=True, y=True):
        """Dataset samples [first, first + count): (x, y, log aspect); x or y
        None when not asked for."""
        xa = np.empty((count, self.nz, self.cin), np.float32) if x else None
        ya = np.empty((count, self.zn, self.cout), np.float32) if y else None
        u = np.empty(count, np.float32)
        check(_lib().hps_nn_data_get(self._h, int(first), int(count),
                                     None if xa is None else _fp(xa),
                                     None if ya is None else _fp(ya), _fp(u)))
        return xa, ya, u

    def fit_norm(self, first=0, count=None, sample=2048, seed=0):
        """out_scale = RMS of the targets over a random sample of the dataset
        range; (u_mu, u_sigma) = mean and spread of its log aspects."""
        n = self.size() - first if count is None else count
        rng = np.random.default_rng(seed)
        ms = []
        for i in np.sort(rng.choice(n, min(n, sample), replace=False)) + first:
            _, y, _ = self.get(int(i), 1, x=False)
            ms.append(float((y.astype(np.float64) ** 2).mean()))
        _, _, u = self.get(first, n, x=False, y=False)
        sig = float(np.std(u))
        self.set_norm(float(np.sqrt(np.mean(ms))), float(np.mean(u)), sig if sig > 1e-3 else 1.0)

    # training and evaluation ---------------------------------------------------
    def train(self, steps, batch=64, lr=1e-3, lr_end=1e-5, seed=0, first=0, count=None):
        """Adam steps on random batches of dataset samples; per-step mean
        relative squared error of the batch."""
        n = self.size() - first if count is None else count
        losses = np.zeros(steps, np.float32)
        check(_lib().hps_nn_train(self._h, int(first), int(n), int(steps), int(batch),
                                  float(lr), float(lr_end), int(seed), _fp(losses)))
        return losses

    def eval(self, first=0, count=None):
        """Relative L2 field error of every sample of a dataset range."""
        n = self.size() - first if count is None else count
        loss = np.zeros(n, np.float32)
        check(_lib().hps_nn_eval(self._h, int(first), int(n), _fp(loss)))
        return np.sqrt(np.maximum(loss, 0.0))

    def predict(self, x, log_aspect):
        x, u = _f32(x), _f32(log_aspect)
        y = np.empty((len(u), self.zn, self.cout), np.float32)
        check(_lib().hps_nn_predict(self._h, _fp(x), _fp(u), len(u), _fp(y)))
        return y

    def loss_grad(self, x, log_aspect, y):
        x, u, y = _f32(x), _f32(log_aspect), _f32(y)
        loss = np.zeros(len(u), np.float32)
        g = np.zeros(self.num_params, np.float32)
        check(_lib().hps_nn_loss_grad(self._h, _fp(x), _fp(u), _fp(y), len(u), _fp(loss), _fp(g)))
        return loss, g


def attach_capture(batch, net):
    """Append every exact-product (charge, field) pair of the batch's kick
    points to net's dataset (None: stop)."""
    check(_lib().hps_batch_set_capture(batch._h, None if net is None else net._h))
    batch._capture = net


def attach(batch, net, audit_every=20, max_err=0.02):
    """Let the network serve the batch's fields, audited against the exact
    product every audit_every kick points with fallback above max_err
    relative L2 error (None: off)."""
    check(_lib().hps_batch_set_surrogate(batch._h, None if net is None else net._h,
                                         int(audit_every), float(max_err)))
    batch._surrogate = net


def info(batch):
    """Counters of the surrogate in a batch."""
    v = np.zeros(6)
    check(_lib().hps_batch_get_surrogate_info(batch._h, N.dptr(v)))
    return {"kick_points": int(v[0]), "audits": int(v[1]), "audit_failures": int(v[2]),
            "max_audit_error": float(v[3]), "network_remeshes": int(v[4]),
            "exact_remeshes": int(v[5])}


# ---------------------------------------------------------------------------
# training data from tracking, training, weights on disk
# ---------------------------------------------------------------------------

def make_batch(bw, members=None, mode="gemm", aspect_step=0.05, sc=True):
    """A Batch of the BatchWorkload's members (all by default) in the given
    field mode; returns (batch, [(beam, spacecharge, simulator)])."""
    from . import HPSim
    from . import configs as cfg
    objs = []
    for k in (range(len(bw.beams)) if members is None else members):
        spec = bw.beams[k]
        b = HPSim.Beam(mass=spec.mass, charge=spec.charge, current=spec.current, num=spec.num)
        b.set_distribution(*cfg.make_distribution(spec))
        b.set_frequency(spec.freq)
        b.set_ref_w(spec.w0)
        s = HPSim.SpaceCharge(nr=bw.nr, nz=bw.nz, interval=bw.interval) if sc else None
        objs.append((b, s, HPSim.Simulator(b, HPSim.BeamLine(elements=bw.lattices[k]), s)))
    batch = HPSim.Batch([m for _, _, m in objs])
    if sc and mode == "gemm":
        batch.set_field_mode("gemm", aspect_step)
    return batch, objs


def training_workloads(nr=32, nz=128, seeds=(101, 102, 103), nbeams=64, ramp_beams=16):
    """Tracking runs whose SCHEFF solves make the training data: batched
    scans of the 40-period lattice (seeded perturbations, 16,384 .. 262,144
    particles) and beams on the 200-period accelerating lattice (0.3 -> 0.6
    MV, its own perturbations); none uses the seeds of the configured
    workloads."""
    from . import configs as cfg
    out = []
    for s in seeds:
        bw = cfg.batched_scan(nbeams=nbeams, nmin=16384, nmax=262144, periods=40, seed=s)
        bw.nr, bw.nz = nr, nz
        out.append(bw)
    base = cfg.config(1, periods=200)
    lat = cfg.build_lattice(base.lattice)
    beams = [cfg.BeamSpec(num=65536, seed=7000 + k) for k in range(ramp_beams)]
    lats = [cfg.perturb_lattice(lat, 9000 + k) for k in range(ramp_beams)]
    out.append(cfg.BatchWorkload(f"ramp-{ramp_beams}x65536x200", beams, lats, nr=nr, nz=nz))
    return out


def stress_workloads(nr=64, nz=256, seeds=(23,), nbeams=16, num=1048576, periods=400):
    """Tracking runs on the combined-stress lattice (the 65,536-particle
    configuration's lattice extended to `periods` periods, unperturbed):
    beams drawn with seeds other than the benchmarked workload's, so the
    training data covers the long, lossy lattice the stress workload
    tracks."""
    from . import configs as cfg
    out = []
    for s in seeds:
        bw = cfg.combined_stress(nbeams=nbeams, num=num, periods=periods, seed=s)
        bw.nr, bw.nz = nr, nz
        out.append(bw)
    return out


def collect(net, workloads, aspect_step=0.05, log=None):
    """Tracks every workload as one grouped batch with the capture on: each
    remesh appends (charge, exact field) to net's dataset."""
    for bw in workloads:
        batch, objs = make_batch(bw, aspect_step=aspect_step)
        attach_capture(batch, net)
        batch.simulate()
        attach_capture(batch, None)
        if log:
            log(f"{bw.name}: dataset {net.size()} samples")
        del batch, objs


def save(net, path, widths, nbot):
    np.savez(path, params=net.params(), norm=np.array(net.norm(), np.float32),
             dims=np.array([net.nr, net.nz, nbot] + list(widths), np.int64))


def load(path, max_batch=64):
    """A Network with the weights and normalisation saved by save()."""
    d = np.load(path, allow_pickle=False)
    nr, nz, nbot = (int(v) for v in d["dims"][:3])
    widths = tuple(int(v) for v in d["dims"][3:7])
    net = Network(nr=nr, nz=nz, widths=widths, nbot=nbot, max_batch=max_batch)
    net.set_params(d["params"])
    net.set_norm(*(float(v) for v in d["norm"]))
    return net


def weights_path(nr, nz):
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                        f"unet_{nr}x{nz}.npz")


def train_network(nr=32, nz=128, widths=DEFAULT_WIDTHS, nbot=DEFAULT_NBOT, steps=20000,
                  batch=64, lr=5e-4, lr_end=5e-6, holdout=0.05, seed=1, workloads=None,
                  capacity=250000, chunk=20000, log=print, init=None, time_limit=None,
                  checkpoint=None):
    """Data from the engine's own SCHEFF solves, then Adam with a cosine
    learning rate from lr to lr_end over all steps, run in chunks of
    `chunk` steps (the held-out error logged after each); returns (net,
    held-out relative L2 errors (a random 5 % of the samples), per-step
    losses).  init: weights file to start from (its normalisation is kept,
    widths and depth come from the file); time_limit: stop after the chunk
    that passes this many seconds of training; checkpoint: file the weights
    are saved to after every chunk."""
    import math
    import time
    if init is not None:
        d = np.load(init, allow_pickle=False)
        nbot = int(d["dims"][2])
        widths = tuple(int(v) for v in d["dims"][3:7])
    net = Network(nr=nr, nz=nz, widths=widths, nbot=nbot, max_batch=batch, seed=seed)
    net.reserve(capacity)
    collect(net, workloads if workloads is not None else training_workloads(nr, nz), log=log)
    net.shuffle(seed)
    n = net.size()
    n_hold = max(1, int(n * holdout))
    if init is not None:
        net.set_params(d["params"])
        net.set_norm(*(float(v) for v in d["norm"]))
        if log:
            e = net.eval(n - n_hold, n_hold)
            log(f"start ({init}): held-out relative L2 error median {np.median(e):.4f}, "
                f"95th percentile {np.percentile(e, 95):.4f}")
    else:
        net.fit_norm(0, n - n_hold)

    def lr_at(s):
        return lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * s / steps))

    parts = []
    t0 = time.time()
    for s0 in range(0, steps, chunk):
        k = min(chunk, steps - s0)
        parts.append(net.train(k, batch=batch, lr=lr_at(s0), lr_end=lr_at(s0 + k),
                               seed=seed + s0, first=0, count=n - n_hold))
        if log:
            e = net.eval(n - n_hold, n_hold)
            log(f"{s0 + k} steps ({time.time() - t0:.0f} s): held-out relative L2 error "
                f"median {np.median(e):.4f}, 95th percentile {np.percentile(e, 95):.4f}")
        if checkpoint:
            save(net, checkpoint, widths, nbot)
        if time_limit is not None and time.time() - t0 > time_limit:
            if log:
                log(f"time limit: stopped after {s0 + k} of {steps} steps")
            break
    losses = np.concatenate(parts)
    err = net.eval(n - n_hold, n_hold)
    if log:
        log(f"trained {net.num_params} parameters on {n - n_hold} samples, {steps} steps: "
            f"held-out relative L2 error median {np.median(err):.4f}, "
            f"95th percentile {np.percentile(err, 95):.4f}, max {err.max():.4f}")
    return net, err, losses


if __name__ == "__main__":
    import argparse
    import os
    ap = argparse.ArgumentParser(description="train the SCHEFF field surrogate on the GPU")
    ap.add_argument("--nr", type=int, default=32)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--steps", type=int, default=200000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--out", default=None)
    ap.add_argument("--init", default=None, help="weights file to start from (fine-tuning)")
    ap.add_argument("--stress-beams", type=int, default=0,
                    help="also capture pairs from this many beams on the stress lattice")
    ap.add_argument("--stress-num", type=int, default=1048576)
    ap.add_argument("--chunk", type=int, default=20000)
    ap.add_argument("--time-limit", type=float, default=None)
    a = ap.parse_args()
    out = a.out or weights_path(a.nr, a.nz)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    wls = training_workloads(a.nr, a.nz)
    if a.stress_beams:
        wls += stress_workloads(a.nr, a.nz, nbeams=a.stress_beams, num=a.stress_num)
    net, err, _ = train_network(a.nr, a.nz, steps=a.steps, lr=a.lr, lr_end=a.lr * 0.01,
                                batch=a.batch, log=lambda s: print(s, flush=True),
                                workloads=wls, init=a.init, chunk=a.chunk,
                                time_limit=a.time_limit, checkpoint=out)
    print(f"saved {out}")
