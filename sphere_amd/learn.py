"""Training of the Darcy correction network on solves the library runs.

The network (sphere_amd.Net, kernels in csrc/net_kernels.hip) maps the
normalized fixed-point residual of a Jacobi iterate, with the cell coupling
and the held-cell mask, to the iterate's normalized error.  Training pairs
come from the engine's own solves: after a fluid step the solve's start and
converged pressures are exported, plain sweeps of the same system give
iterates p_k, and each pair is (input(p_k), (p_final - p_k) / s_k).

PyTorch supplies the device memory, the optimizer (Adam on the FP32 master
parameters) and the loss; the network's forward and backward passes are the
library's kernels.
"""
import numpy

from ._abi import Net

SWEEPS = (2, 12, 22, 42, 82, 162, 322)

# layer shapes (cin, cout, taps) of the network, in parameter order
LAYERS = [(3, 16, 27), (16, 32, 27), (32, 48, 27), (48, 64, 27), (64, 48, 27), (48, 32, 27),
          (32, 16, 27), (16, 16, 27), (16, 1, 1)]


def init_params(seed=0):
    """He-normal weights (std sqrt(2 / (taps cin))), zero biases; the output
    layer starts small so that an untrained network changes little."""
    rng = numpy.random.default_rng(seed)
    out = []
    for k, (ci, co, taps) in enumerate(LAYERS):
        std = numpy.sqrt(2.0 / (taps * ci)) * (0.1 if k == len(LAYERS) - 1 else 1.0)
        out.append(rng.standard_normal(taps * co * ci) * std)
        out.append(numpy.zeros(co))
    p = numpy.concatenate(out).astype(numpy.float32)
    assert p.size == Net.param_count()
    return p


def grid_of(engine):
    h = engine.header()
    return int(h.num[0]), int(h.num[1]), int(h.num[2])


def periodic_of(engine):
    """Network wrap bits of the Darcy boundary conditions (bit 0 x, bit 1 y),
    as the library chooses them for its own solves."""
    h = engine.header()
    return (1 if h.bc_xn == 2 and h.bc_xp == 2 else 0) | (2 if h.bc_yn == 2 and h.bc_yp == 2 else 0)


def make_net(engine):
    nx, ny, nz = grid_of(engine)
    return Net(nx, ny, nz, periodic=periodic_of(engine))


def correct(engine, net, params_t, p, inp, out):
    """The library's learned correction of iterate p (in place): input,
    network, p += s e on the cells that are not held, ghosts refreshed."""
    import torch
    nx, ny, nz = grid_of(engine)
    s = engine.darcy_net_input(p, inp)
    if not (0.0 < s < 1e300):
        return
    net.forward(params_t, inp, out)
    keep = (inp[:, 2] < 0.5).to(torch.float64)
    pin = p.view(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1]
    pin += (s * out.to(torch.float64) * keep).view(nz, ny, nx)
    engine.darcy_sweeps(p, 0)  # ghost layer of the corrected iterate


class PairStore:
    """Training pairs in preallocated device tensors: inputs [N, V, 3] in
    BF16 (the network rounds its input to BF16 as its first operation, so
    this is what it sees) and targets [N, V] in FP32."""

    def __init__(self, capacity, V):
        import torch
        dev = torch.device("cuda:0")
        self.X = torch.empty((capacity, V, 3), dtype=torch.bfloat16, device=dev)
        self.Y = torch.empty((capacity, V), dtype=torch.float32, device=dev)
        self.n = 0

    def full(self):
        return self.n >= self.X.shape[0]

    def add(self, inp, target):
        if self.full():
            return False
        self.X[self.n].copy_(inp)
        self.Y[self.n].copy_(target)
        self.n += 1
        return True

    def view(self):
        return self.X[:self.n], self.Y[:self.n]


def collect(engine, steps, sweeps=SWEEPS, log=None, net=None, params=None, corrections=8,
            store=None):
    """Run `steps` fluid steps of the engine's loaded state; after each,
    make training pairs for the iterates after each count in `sweeps` of
    plain sweeps from the solve's start.  With a network and its parameters,
    also for the iterates the corrected solve meets: 2 sweeps, then
    (correction, 10 sweeps) repeated, a pair before each correction.
    Returns (inputs [S, V, 3], targets [S, V]) as float32 CUDA tensors, or,
    with a PairStore, adds the pairs to it (until it is full) and returns
    it."""
    import torch
    nx, ny, nz = grid_of(engine)
    V = nx * ny * nz
    cells = (nx + 2) * (ny + 2) * (nz + 2)
    dev = torch.device("cuda:0")
    p0 = torch.empty(cells, dtype=torch.float64, device=dev)
    pf = torch.empty(cells, dtype=torch.float64, device=dev)
    inputs, targets = [], []
    order = sorted(sweeps)
    for k in range(steps):
        if store is not None and store.full():
            break
        engine.step(1)
        st = engine.darcy_stats()
        engine.darcy_export(0, p0)
        engine.darcy_export(1, pf)
        pfi = pf.view(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1].reshape(-1)
        pk = p0.clone()
        done = 0
        def pair(p):
            inp = torch.empty((V, 3), dtype=torch.float32, device=dev)
            s = engine.darcy_net_input(p, inp)
            if not (0.0 < s < 1e300):
                return
            pi = p.view(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1].reshape(-1)
            if store is not None:
                store.add(inp, ((pfi - pi) / s).to(torch.float32))
                return
            inputs.append(inp)
            targets.append(((pfi - pi) / s).to(torch.float32))

        for n in order:
            if n > st.last_nijac:
                break
            engine.darcy_sweeps(pk, n - done)
            done = n
            pair(pk)
        if net is not None and params is not None:
            pt = torch.tensor(params, dtype=torch.float32, device=dev)
            inp = torch.empty((V, 3), dtype=torch.float32, device=dev)
            out = torch.empty(V, dtype=torch.float32, device=dev)
            pc = p0.clone()
            engine.darcy_sweeps(pc, 2)
            for c in range(corrections):
                if 2 + 10 * c > st.last_nijac:
                    break
                pair(pc)
                correct(engine, net, pt, pc, inp, out)
                engine.darcy_sweeps(pc, 10)
        if log and (k + 1) % 50 == 0:
            log("collected %d pairs after %d fluid steps" %
                (store.n if store is not None else len(inputs), k + 1))
    if store is not None:
        return store
    return torch.stack(inputs), torch.stack(targets)


def train(net, inputs, targets, iters=2000, lr=1e-3, seed=0, params=None, log=None):
    """Adam on the mean squared error of the normalized error estimate.
    Returns the FP32 parameters (numpy) and the loss history."""
    import torch
    dev = inputs.device
    p = torch.tensor(init_params(seed) if params is None else params, dtype=torch.float32,
                     device=dev).requires_grad_(True)
    opt = torch.optim.Adam([p], lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, iters, eta_min=lr * 0.05)
    V = inputs.shape[1]
    out = torch.empty(V, dtype=torch.float32, device=dev)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    for it in range(iters):
        i = int(torch.randint(inputs.shape[0], (1,), generator=gen))
        with torch.no_grad():
            net.forward(p, inputs[i].float(), out)
            diff = out - targets[i]
            loss = torch.mean(diff * diff)
            gout = (2.0 / V) * diff
            g = torch.zeros_like(p)
            net.backward(p, gout, g)
        p.grad = g
        opt.step()
        sched.step()
        if it % 50 == 0 or it == iters - 1:
            hist.append((it, float(loss)))
            if log and it % 500 == 0:
                log("train iter %d loss %.3e" % (it, float(loss)))
    return p.detach().cpu().numpy(), hist


def evaluate(net, params, inputs, targets):
    """Mean squared error and the mean relative error norm of the estimates
    (||e_hat - e|| / ||e||) over the pairs."""
    import torch
    dev = inputs.device
    p = torch.tensor(params, dtype=torch.float32, device=dev)
    out = torch.empty(inputs.shape[1], dtype=torch.float32, device=dev)
    mse, rel = [], []
    with torch.no_grad():
        for i in range(inputs.shape[0]):
            net.forward(p, inputs[i].float(), out)
            d = out - targets[i]
            mse.append(float(torch.mean(d * d)))
            rel.append(float(torch.linalg.norm(d) / torch.linalg.norm(targets[i]).clamp_min(1e-30)))
    return float(numpy.mean(mse)), float(numpy.mean(rel))
