"""Learned warm start: a small network that maps the single-scattering source of a cell (and the optical properties
there) to the converged multiple-scattering source, used only as the starting point x0 of Solver.solve; the source
iteration then runs to its tolerance as without it. No counterpart in the reference.

Network: a multilayer perceptron, 4 linear layers of width 256 with GELU between them, BF16 weights (FP32 master
copies for Adam) and FP32 accumulation; every product, the loss, the GELU derivative, the bias sums and Adam run in
the HIP kernels of libsosfuse2_hip.so (sos_mlp_*); PyTorch holds the buffers and assembles the features.

Samples are (cell, lane) pairs. Inputs: the cell's first-light source rows (tensor rows, SH x Rayleigh rows and the
SH planes) divided by the lane's scale s = max over cells of the first-light tensor trace, and 16 physical features
(height, cos of the solar-zenith angle, log local extinction of the six components, the lane's asymmetries and
albedo, log vertical optical depths of Rayleigh, aerosol, absorbers and sulfate, log s). Target: the multiple-
scattering increment (converged - first-light source) / s. Inputs and targets are standardised per row with
statistics of the training set (kept with the network), so that small rows (the xz tensor row, high SH orders)
weigh in the loss as much as large ones. The warm start is x0 = first + increment in sunlit cells, with the tensor
diagonal kept at or above the first-light value (multiple scattering only adds light), and 0 in shadow cells.
"""

import math

import numpy as np
import torch

from . import atmosphere, hip

N_PHYS = 16


class Network:
    """MLP sizes[0] -> sizes[1] -> ... -> sizes[-1] on the HIP kernels; parameters in one flat FP32 buffer."""

    def __init__(self, sizes, device='cuda', seed=0):
        self.sizes = list(sizes)
        self.device = torch.device(device)
        rng = np.random.default_rng(seed)
        shapes = []
        for i, o in zip(sizes, sizes[1:]):
            shapes += [(o, i), (o,)]
        n = sum(int(np.prod(s)) for s in shapes)
        flat = np.zeros(n, 'f4')
        self.views = []
        off = 0
        for k, s in enumerate(shapes):
            size = int(np.prod(s))
            if k % 2 == 0:  # weight: scaled for GELU layers (He-like), biases start at zero
                flat[off : off + size] = rng.normal(size=size) * math.sqrt(2.0 / s[1])
            self.views.append((off, s))
            off += size
        self.flat = torch.from_numpy(flat).to(self.device)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.steps = 0
        self.in_mu = self.in_sd = self.out_sd = None  # per-row standardisation of inputs and targets
        self._work = None
        self._lossbuf = torch.zeros(1, dtype=torch.float32, device=self.device)
        self._lwork = torch.zeros(1024, dtype=torch.float64, device=self.device)

    @property
    def layers(self):
        return len(self.sizes) - 1

    def _param(self, buf, k):
        off, s = self.views[k]
        return buf[off : off + int(np.prod(s))].view(*s)

    def W(self, i, buf=None):
        return self._param(self.flat if buf is None else buf, 2 * i)

    def b(self, i, buf=None):
        return self._param(self.flat if buf is None else buf, 2 * i + 1)

    def state_dict(self):
        return dict(sizes=self.sizes, flat=self.flat.cpu().numpy(), steps=self.steps, in_mu=self.in_mu.cpu().numpy(),
                    in_sd=self.in_sd.cpu().numpy(), out_sd=self.out_sd.cpu().numpy())

    def load(self, d):
        assert list(d['sizes']) == self.sizes
        self.flat.copy_(torch.from_numpy(np.asarray(d['flat'], 'f4')))
        for k in ('in_mu', 'in_sd', 'out_sd'):
            setattr(self, k, torch.from_numpy(np.asarray(d[k], 'f4')).to(self.device))

    def fit_normalization(self, X, T):
        """Per-row mean / deviation of the inputs and deviation of the targets over a training set."""
        self.in_mu = X.mean(0)
        self.in_sd = X.std(0).clamp_min(1e-6)
        self.out_sd = T.pow(2).mean(0).sqrt().clamp_min(1e-6)

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _work_for(self, n):
        if self._work is None or self._work.numel() < n:
            self._work = torch.empty(n, dtype=torch.float32, device=self.device)
        return self._work

    # ----------------------------------------------------------------------------------------------- passes
    def forward(self, x, keep=False):
        """y [N, sizes[-1]] for x [N, sizes[0]] (FP32, contiguous). keep=True keeps the layer inputs and
        pre-activations for backward."""
        N = x.shape[0]
        h = x
        self._cache = [] if keep else None
        for i in range(self.layers):
            o, n_in = self.sizes[i + 1], self.sizes[i]
            z = torch.empty((N, o), dtype=torch.float32, device=self.device)
            last = i == self.layers - 1
            hn = None if last else torch.empty((N, o), dtype=torch.float32, device=self.device)
            hip.call('sos_mlp_gemm', N, o, n_in, hip.ptr(h), h.stride(0), 1, hip.ptr(self.W(i)), 1, n_in,
                     hip.ptr(self.b(i)), 0 if last else 1, hip.ptr(z), o, hip.ptr(hn), o, 1, None, self._stream())
            if keep:
                self._cache.append((h, z))
            h = z if last else hn
        return h

    def loss_and_grad(self, y, t, weight=None):
        """Mean weighted squared error and dL/dy."""
        dy = torch.empty_like(y)
        hip.call('sos_mlp_mse', y.shape[0], y.shape[1], hip.ptr(y), y.stride(0), hip.ptr(t), t.stride(0),
                 hip.ptr(weight), hip.ptr(dy), dy.stride(0), hip.ptr(self._lossbuf), hip.ptr(self._lwork),
                 self._stream())
        return self._lossbuf, dy

    def loss(self, y, t, weight=None):
        hip.call('sos_mlp_mse', y.shape[0], y.shape[1], hip.ptr(y), y.stride(0), hip.ptr(t), t.stride(0),
                 hip.ptr(weight), None, 0, hip.ptr(self._lossbuf), hip.ptr(self._lwork), self._stream())
        return self._lossbuf

    def backward(self, dy, splits=8):
        """Parameter gradients into self.grad from dL/dy of the last forward(keep=True)."""
        N = dy.shape[0]
        dz = dy
        for i in range(self.layers - 1, -1, -1):
            o, n_in = self.sizes[i + 1], self.sizes[i]
            h, _ = self._cache[i]
            sp = max(1, min(splits, N // 256))
            work = self._work_for(sp * o * n_in) if sp > 1 else None
            # dW = dz^T h: M = o, N = n_in, K = samples
            hip.call('sos_mlp_gemm', o, n_in, N, hip.ptr(dz), 1, dz.stride(0), hip.ptr(h), h.stride(0), 1, None, 0,
                     hip.ptr(self.W(i, self.grad)), n_in, None, 0, sp, hip.ptr(work), self._stream())
            hip.call('sos_mlp_bias_grad', N, o, hip.ptr(dz), dz.stride(0), hip.ptr(self.b(i, self.grad)),
                     self._stream())
            if i:
                dh = torch.empty((N, n_in), dtype=torch.float32, device=self.device)
                hip.call('sos_mlp_gemm', N, n_in, o, hip.ptr(dz), dz.stride(0), 1, hip.ptr(self.W(i)), n_in, 1, None,
                         0, hip.ptr(dh), n_in, None, 0, 1, None, self._stream())
                _, zprev = self._cache[i - 1]
                hip.call('sos_mlp_gelu_backward', N, n_in, hip.ptr(dh), n_in, hip.ptr(zprev), zprev.stride(0),
                         hip.ptr(dh), n_in, self._stream())
                dz = dh
        self._cache = None

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.steps += 1
        hip.call('sos_mlp_adam', self.flat.numel(), hip.ptr(self.flat), hip.ptr(self.grad), hip.ptr(self.m),
                 hip.ptr(self.v), self.steps, lr, beta1, beta2, eps, self._stream())

    def predict(self, x, batch=65536):
        """Targets in physical units for raw inputs (the standardisation is applied and undone here)."""
        out = torch.empty((x.shape[0], self.sizes[-1]), dtype=torch.float32, device=self.device)
        for i in range(0, x.shape[0], batch):
            xb = ((x[i : i + batch] - self.in_mu) / self.in_sd).contiguous()
            out[i : i + batch] = self.forward(xb) * self.out_sd
        return out


# ------------------------------------------------------------------------------------------------- features
def first_source(solver):
    """The first-light source [cells, rows, G]: the composition of the first-light moments alone."""
    s, k = solver.s, solver.kernel
    solver.initialize()
    dc = k.constants(s, solver.geometry.bundle.nq)
    out = torch.empty((s.geo.ns, s.rows, dc.G), dtype=torch.float32, device=k.device)
    hip.call('sos_compose', dc.struct, s.rows, hip.ptr(solver.first_moments), None, hip.ptr(dc.factor), hip.ptr(out),
             k.stream())
    return out


def _lane_physics(s):
    """Per-lane [W, 7] and per-(cell, lane) [cells, W, 8] physical features (host FP64, small)."""
    o, g, st = s.optics, s.geo, s.state
    h = g.h
    prof = atmosphere.profiles(st, h)                       # [nh, 5]
    sprof = atmosphere.strato_profile(st, h)                # [nh]
    dh = np.gradient(h)
    col = (prof * dh[:, None]).sum(0)                       # vertical columns of the unit profiles
    scol = float((sprof * dh).sum())
    tau = np.concatenate([o.ext * col[:, None], (o.strato_ext * scol)[None]], 0)   # [6, W]
    lane = np.stack([o.g, o.strato_g, o.albedo, np.log10(np.maximum(tau[0], 1e-12)),
                     np.log10(np.maximum(tau[1], 1e-12)), np.log10(np.maximum(tau[2] + tau[3] + tau[4], 1e-12)),
                     np.log10(np.maximum(tau[5], 1e-12))], -1)                           # [W, 7]
    local = np.concatenate([prof[:, :, None] * o.ext[None], (sprof[:, None] * o.strato_ext[None])[:, None]], 1)
    loc = np.log10(np.maximum(local, 1e-12))                                            # [nh, 6, W]
    ih = np.repeat(np.arange(g.nh), g.nc)
    chi = np.tile(g.chi, g.nh)
    cell = np.concatenate([(h[ih] / g.domain.top_km)[:, None, None].repeat(s.W, 1),
                           np.cos(chi)[:, None, None].repeat(s.W, 1),
                           np.transpose(loc[ih], (0, 2, 1))], -1)                        # [cells, W, 8]
    return lane, cell


def features(s, first, lanes=None, cells=None):
    """Network inputs [n, rows + 16] for the (cell, lane) pairs of `cells` x `lanes`, and the lane scales."""
    W = s.W
    lanes = np.arange(W) if lanes is None else np.asarray(lanes)
    trace = first[:, 0:3, :W].sum(1)                                         # [cells, W]
    scale = torch.clamp(trace.abs().amax(0), min=1e-30)                      # [W]
    lane_f, cell_f = _lane_physics(s)
    dev = first.device
    cells = torch.arange(s.geo.ns, device=dev) if cells is None else torch.as_tensor(cells, device=dev)
    out = []
    for w in lanes:
        w = int(w)
        src = first[cells, :, w] / scale[w]
        phys = np.concatenate([cell_f[:, w], np.broadcast_to(lane_f[w], (s.geo.ns, 7)),
                               np.full((s.geo.ns, 1), math.log10(float(scale[w])))], -1)
        ph = torch.from_numpy(np.ascontiguousarray(phys, 'f4')).to(dev)[cells]
        out.append(torch.cat([src, ph], 1))
    return torch.cat(out, 0).contiguous(), scale


def sizes_for(rows, width=256):
    return [rows + N_PHYS, width, width, width, rows]


def samples(s, solver, x, cells_per_lane=None, rng=None):
    """Training pairs (inputs, targets) of one solved context: x the converged source; the targets are the
    increments (x - first) / s."""
    first = first_source(solver)
    C = s.geo.ns
    if cells_per_lane is not None and cells_per_lane < C:
        cells = np.sort((rng or np.random.default_rng(0)).choice(C, cells_per_lane, replace=False))
    else:
        cells = np.arange(C)
    X, scale = features(s, first, cells=cells)
    cd = torch.as_tensor(cells, device=x.device)
    T = torch.cat([(x[cd, :, w] - first[cd, :, w]) / scale[w] for w in range(s.W)], 0).contiguous()
    return X, T


def train(net, X, T, steps, batch=4096, lr=1e-3, seed=0, log=None):
    """Adam on random minibatches of the standardised data; returns the loss history (one value per 50 steps)."""
    if net.in_mu is None:
        net.fit_normalization(X, T)
    g = torch.Generator(device='cpu').manual_seed(seed)
    n = X.shape[0]
    hist = []
    for step in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), generator=g).to(X.device)
        xb = ((X[idx] - net.in_mu) / net.in_sd).contiguous()
        tb = (T[idx] / net.out_sd).contiguous()
        y = net.forward(xb, keep=True)
        loss, dy = net.loss_and_grad(y, tb)
        net.backward(dy)
        net.adam(lr)
        if step % 50 == 0 or step == steps - 1:
            hist.append(float(loss.item()))
            if log:
                log(f'LEARN step {step} loss {hist[-1]:.4e}')
    return hist


def warm_start(net, s, solver, lit=1e-3, rows='all'):
    """x0 [cells, rows, G] for Solver.solve: the first-light source, plus in sunlit cells (first-light trace >= lit x
    the lane's largest) the network's multiple-scattering increment, with the tensor diagonal kept at or above its
    first-light value. Shadow cells keep the first-light source because their traces are tiny: the log-space corner
    blend magnifies a relative error there, and a guess the iteration must undo costs more evaluations than it saves.
    rows='low' takes the increment only for the tensor rows and the isotropic (l = 0) SH rows and keeps the
    first-light values of the higher orders, whose predictions are the least accurate. Lanes beyond W repeat the
    last lane (the padding rule of the lane table)."""
    first = first_source(solver)
    X, scale = features(s, first)
    Y = net.predict(X)
    C, W = s.geo.ns, s.W
    if rows == 'low':
        keep = torch.zeros(first.shape[1], dtype=torch.bool, device=first.device)
        keep[0:11] = True
        if first.shape[1] > 10 + s.K:
            keep[10 + s.K] = True
        Y = Y * keep[None, :].to(Y.dtype)
    x0 = torch.empty_like(first)
    for w in range(first.shape[2]):
        wl = min(w, W - 1)
        x0[:, :, w] = first[:, :, wl] + Y[wl * C : (wl + 1) * C] * scale[wl]
    x0[:, 0:3] = torch.maximum(x0[:, 0:3], first[:, 0:3])
    trace = first[:, 0:3].sum(1)
    sunlit = trace.abs() >= lit * trace.abs().amax(0, keepdim=True)
    return torch.where(sunlit[:, None, :], x0, first)
