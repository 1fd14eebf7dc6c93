"""The learned SCHEFF field surrogate: network wrapper, training data from the
engine's own SCHEFF solve, and the normalisation that makes the field
dimensionless.

The SCHEFF field table is linear in the charge histogram q and, for a fixed
mesh shape, scales as c1 / dr^2 (the Green's table entries, space_charge_kernel.cu:38-100,
are c1 times ring fields with lengths in metres).  So

    E(zp, rp) = (c1 Q / dr^2) * F(q / Q, dz gamma / dr)

with Q the total charge on the mesh and F dimensionless.  The network learns
F / s (s: the RMS of F over its training set, kept with the network as its
output scale) on the (nz+1) x (nr+1) field mesh from four input channels:
nr nz q / Q (bins at rows zs < nz, columns rs < nr; zero elsewhere; mean 1
over the mesh), log(dz gamma / dr), zp / nz - 0.5 and rp / nr.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _native as N
from ._native import check


def _lib():
    return N.lib()


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Network:
    """The surrogate network (hps_nn_*): H x W field mesh, `width` channels."""

    def __init__(self, H, W, width=32, max_batch=32, seed=1):
        h = C.c_void_p()
        check(_lib().hps_nn_create(C.byref(h), int(H), int(W), int(width), int(max_batch), int(seed)))
        self._h = h
        self.H, self.W, self.width, self.max_batch = int(H), int(W), int(width), int(max_batch)

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
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        check(_lib().hps_nn_set_params(self._h, _fp(p)))

    def layers(self):
        n = C.c_uint32()
        check(_lib().hps_nn_layers(self._h, None, 0, C.byref(n)))
        t = np.zeros((n.value, 8), np.int32)
        check(_lib().hps_nn_layers(self._h, t.ctypes.data_as(C.POINTER(C.c_int32)), n.value,
                                   C.byref(n)))
        return t

    @property
    def out_scale(self):
        v = C.c_float()
        check(_lib().hps_nn_get_output_scale(self._h, C.byref(v)))
        return v.value

    @out_scale.setter
    def out_scale(self, v):
        check(_lib().hps_nn_set_output_scale(self._h, float(v)))

    def set_data(self, x, y):
        """Training set: inputs x and dimensionless fields y; the output scale
        becomes the RMS of y and the network is trained on y / scale."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.asarray(y, np.float32)
        assert x.shape[1:] == (self.H, self.W, 4) and y.shape[1:] == (self.H, self.W, 2)
        s = float(np.sqrt(np.mean(np.square(y, dtype=np.float64))))
        self.out_scale = s
        ys = np.ascontiguousarray(y / s, np.float32)
        check(_lib().hps_nn_set_data(self._h, _fp(x), _fp(ys), x.shape[0]))

    def train(self, steps, batch=16, lr=1e-3, seed=1):
        losses = np.empty(int(steps), np.float32)
        check(_lib().hps_nn_train(self._h, int(steps), int(batch), float(lr), int(seed), _fp(losses)))
        return losses

    def predict(self, x):
        """The dimensionless field F (the network output times its scale)."""
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty(x.shape[:3] + (2,), np.float32)
        check(_lib().hps_nn_predict(self._h, _fp(x), x.shape[0], _fp(out)))
        return out

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = np.empty(x.shape[0], np.float32)
        g = np.empty(self.num_params, np.float32)
        check(_lib().hps_nn_loss_grad(self._h, _fp(x), _fp(y), x.shape[0], _fp(loss), _fp(g)))
        return loss, g


def field_scale(state, beam_current, beam_freq, mass, num, frac=1.0):
    """c1 / dr^2 of the SCHEFF tables of one remesh: the factor between the
    dimensionless field F and the field table (per unit of total charge)."""
    c1 = 5721.67 * (beam_current * frac / mass) / (beam_freq * 1e6 * num)
    return c1 / (state["dr"] * state["dr"])


def make_input(bins, nr, nz, aspect):
    """Input channels [nz+1][nr+1][4] of one histogram (int64 bins [nz][nr] in
    units of 2^-36); returns (x, Q) with Q the total charge in particles."""
    q = np.asarray(bins, np.float64).reshape(nz, nr) * 2.0 ** -36
    Q = float(q.sum())
    x = np.zeros((nz + 1, nr + 1, 4), np.float32)
    if Q > 0:
        x[:nz, :nr, 0] = q * (nr * nz / Q)
    x[..., 1] = math.log(aspect)
    x[..., 2] = (np.arange(nz + 1, dtype=np.float32) / nz - 0.5)[:, None]
    x[..., 3] = (np.arange(nr + 1, dtype=np.float32) / nr)[None, :]
    return x, Q


def aspect_of(state, gamma):
    return state["dz"] * gamma / state["dr"]


def rel_l2(a, b):
    """Relative L2 error of a against b."""
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def collect_pairs(beam_spec, elements, nr=32, nz=128, interval=0.05, max_pairs=None):
    """Track one beam element by element on the GPU and keep (input, F,
    scale) at every SCHEFF remesh: input [nz+1][nr+1][4], the dimensionless
    field F = field table / (c1 Q / dr^2) [nz+1][nr+1][2] and that factor."""
    from . import HPSim
    from . import configs as cfg
    b = HPSim.Beam(mass=beam_spec.mass, charge=beam_spec.charge, current=beam_spec.current,
                   num=beam_spec.num)
    b.set_distribution(*cfg.make_distribution(beam_spec))
    b.set_frequency(beam_spec.freq)
    b.set_ref_w(beam_spec.w0)
    sc = HPSim.SpaceCharge(nr=nr, nz=nz, interval=interval)
    sim = HPSim.Simulator(b, HPSim.BeamLine(elements=elements), sc)
    sim.set_graph(False)
    xs, fs, scales = [], [], []
    last = 0
    for i in range(len(elements)):
        sim.simulate(i, i)
        st = sc.state()
        if st["remesh_count"] == last:
            continue
        last = st["remesh_count"]
        _, fld2, bins = sc.tables()
        gamma = b.stats()["w_avg_good"] / beam_spec.mass + 1.0
        x, Q = make_input(bins, nr, nz, aspect_of(st, gamma))
        if Q <= 0:
            continue
        scale = field_scale(st, b.get_current(), b.get_frequency(), beam_spec.mass, beam_spec.num) * Q
        xs.append(x)
        fs.append((fld2.reshape(nz + 1, nr + 1, 2) / scale).astype(np.float32))
        scales.append(scale)
        if max_pairs and len(xs) >= max_pairs:
            break
    return np.array(xs), np.array(fs), np.array(scales)


def training_set(runs, num=65536, periods=40, seed=1, nr=32, nz=128):
    """(x, F) pairs from `runs` tracking runs of the 40-period lattice with
    varied current, transverse and longitudinal beam sizes (seeded)."""
    from . import configs as cfg
    rng = np.random.default_rng(seed)
    lattice = cfg.build_lattice(cfg.config(0, periods=periods).lattice)
    X, F = [], []
    for k in range(runs):
        spec = cfg.BeamSpec(num=num, seed=seed * 1000 + k,
                            current=float(rng.uniform(0.002, 0.03)),
                            sig_x=1e-3 * float(rng.uniform(0.6, 1.6)),
                            sig_y=1e-3 * float(rng.uniform(0.6, 1.6)),
                            sig_phi=math.radians(float(rng.uniform(5.0, 20.0))),
                            sig_w=2e-3 * float(rng.uniform(0.5, 2.0)))
        x, f, _ = collect_pairs(spec, lattice, nr=nr, nz=nz)
        X.append(x)
        F.append(f)
    return np.concatenate(X), np.concatenate(F)
