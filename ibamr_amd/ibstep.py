"""Several Lagrangian structures on one level, each with its own delta kernel,
stepped force -> spread -> velocity -> ghost fill -> interp -> position update,
eagerly or as one captured HIP graph.

Reference behaviour mirrored:
  * IBStrategySet runs each strategy's spreadForce in turn into the same f, so
    the spreads accumulate, and each strategy interpolates its own points
    (src/IB/IBStrategySet.cpp:162-172, 215-225).  Here every contiguous run of
    structures sharing a kernel is one level plan; the first plan overwrites f,
    the others accumulate.
  * IBMethod::forwardEulerStep: X_new = X + dt U (src/IB/IBMethod.cpp:713-739).
  * Positions that leave the periodic box are moved back into it every step,
    as LDataManager::beginDataRedistribution does at a regrid (X += L while
    X < x_lower, then X -= L while X >= x_upper, LDataManager.cpp:1403-1407):
    the points are binned every step here, and a point outside the domain would
    otherwise drop out of the spread and keep a stale velocity.
  * With a force generator (ibamr_amd.structure.Forces), F is set to the force
    at X and the previous U at the start of every step, as
    IBMethod::computeLagrangianForce zeroes F and calls the force generator
    before the spread (src/IB/IBMethod.cpp:834-857).

The fluid solve between spread and interp is not part of this library; the
step uses a stand-in u = u_bg + alpha * f on the whole ghosted arrays, followed
by the periodic ghost fill, so the interpolated velocity depends on the spread.

Graph capture (torch.cuda.CUDAGraph) records only kernels: the force
generation, point binning (ibamr_plan_bin_async), spread, the stand-in
combination (a torch elementwise kernel), the periodic ghost fill, interp, and
the Euler update with the periodic wrap (one library kernel).  No memset and
no host synchronization happens inside a step.
"""
import torch

import ibamr_amd as ia


class StructureSet:
    """structures: list of (kernel, npoints) in the order their points sit in X.
    forces: None (F is given) or a structure.Forces over all the points."""

    def __init__(self, level, structures, precision=ia.FP64, alpha=0.0, u_bg=None, device="cuda", forces=None):
        self.level = level
        self.forces = forces
        self.precision = precision
        self.dtype = torch.float64 if precision == ia.FP64 else torch.float32
        self.alpha = float(alpha)
        gw = max(ia.get_minimum_ghost_width(k) for k, _ in structures)
        self.gcw = [gw] * level.ndim
        # contiguous runs of one kernel -> one plan each
        self.groups = []
        start = 0
        for k, n in structures:
            if self.groups and self.groups[-1][0] == k:
                self.groups[-1][2] += n
            else:
                self.groups.append([k, start, n])
            start += n
        self.npoints = start
        self.plans = [ia.LevelPlan(level, k, ia.SIDE, precision, data_gcw=self.gcw, max_points=max(n, 1))
                      for k, _, n in self.groups]
        shp = [ia.array_shape(level, 0, self.gcw, ia.SIDE, axis=a) for a in range(level.ndim)]
        self.f = [torch.zeros(s, dtype=self.dtype, device=device) for s in shp]
        self.u = [torch.zeros(s, dtype=self.dtype, device=device) for s in shp]
        self.u_bg = [torch.zeros(s, dtype=self.dtype, device=device) for s in shp] if u_bg is None else \
            [b.to(self.dtype).contiguous() for b in u_bg]
        for p in self.plans:
            p.bind(0, self.f)
            p.bind(1, self.u)
        self.graph = None

    def _slices(self):
        return [(p, s, s + n) for p, (_, s, n) in zip(self.plans, self.groups)]

    def step(self, X, F, U, dt, stream=None):
        """One step on device tensors X [n][ndim] float64, F/U [n][ndim] in the
        plan precision; X is advanced in place."""
        s = stream if stream is not None else torch.cuda.current_stream()
        if self.forces is not None:
            self.forces.compute(F, X, U, accumulate=False, stream=s)
        for p, a, b in self._slices():
            p.bin(X[a:b], stream=s, check=False)
        for i, (p, a, b) in enumerate(self._slices()):
            p.spread(0, F[a:b], X[a:b], mode=ia.OVERWRITE if i == 0 else ia.ACCUMULATE, stream=s)
        with torch.cuda.stream(s):
            for d in range(self.level.ndim):
                torch.add(self.u_bg[d], self.f[d], alpha=self.alpha, out=self.u[d])
        self.plans[0].fill_periodic_ghosts(1, stream=s)
        # X has not moved since the spread: each plan interpolates from the
        # bin-ordered copy (and, FP32C, the stencils) its spread set up
        for p, a, b in self._slices():
            p.interp(1, U[a:b], None, stream=s)
        # X + dt U and the move back into the periodic box in one kernel (one
        # period suffices: a step moves a point by less than one period)
        ia.forward_euler_step_periodic(X, X, U, dt, self.level, stream=s)

    def autotune(self, X, F, reps=2, stream=None):
        """Picks each plan's spread tile shape for its structures' current
        positions (LevelPlan.autotune: trial spreads overwrite the force arrays,
        which the next step overwrites anyway).  Returns {kernel: {shape: ms}}.
        A graph captured before keeps replaying the shapes it captured (every
        shape's work list lives as long as its plan); capture again to use the
        chosen ones."""
        out = {}
        for (k, _, _), (p, a, b) in zip(self.groups, self._slices()):
            p.bin(X[a:b], stream=stream, check=False)
            out[k] = p.autotune(0, F[a:b], X[a:b], reps=reps, stream=stream)
        return out

    def capture(self, X, F, U, dt, steps_per_graph=1):
        """Captures steps_per_graph steps into one HIP graph (after one eager
        warm-up step has been run by the caller)."""
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                for _ in range(steps_per_graph):
                    self.step(X, F, U, dt, stream=side)
        torch.cuda.current_stream().wait_stream(side)
        self.graph = g
        return g


def ellipsoid_shells(nshell, counts, seed, centre_range=(0.2, 0.8), axes_range=(0.05, 0.15), device="cuda"):
    """Points on nshell ellipsoids (centres uniform in centre_range^3, semi-axes
    uniform in axes_range, directions uniform), counts[i] points on shell i,
    wrapped into the periodic unit cube.  Returns float64 [sum(counts)][3]."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    out = []
    lo, hi = centre_range
    alo, ahi = axes_range
    for i in range(nshell):
        c = lo + (hi - lo) * torch.rand(3, generator=g, device=device, dtype=torch.float64)
        ax = alo + (ahi - alo) * torch.rand(3, generator=g, device=device, dtype=torch.float64)
        d = torch.randn((counts[i], 3), generator=g, device=device, dtype=torch.float64)
        d = d / d.norm(dim=1, keepdim=True)
        out.append(c + d * ax)
    return torch.remainder(torch.cat(out), 1.0).contiguous()


def background_velocity(level, gcw, dtype, amplitude=0.1, device="cuda"):
    """A periodic divergence-free face field: u0 = A sin(2 pi y) cos(2 pi z),
    u1 = A sin(2 pi z) cos(2 pi x), u2 = A sin(2 pi x) cos(2 pi y), at the faces
    of the ghosted side arrays (periodic, so ghosts hold the periodic values)."""
    import math
    p = level.patches[0]
    h = level.dx[0]
    out = []
    for a in range(3):
        n = [p.iupper[d] - p.ilower[d] + 1 + 2 * gcw[d] + (1 if d == a else 0) for d in range(3)]
        c = [(torch.arange(n[d], device=device, dtype=torch.float64) - gcw[d] + (0.0 if d == a else 0.5)) * h
             for d in range(3)]
        z, y, x = torch.meshgrid(c[2], c[1], c[0], indexing="ij")
        t = 2 * math.pi
        if a == 0:
            v = torch.sin(t * y) * torch.cos(t * z)
        elif a == 1:
            v = torch.sin(t * z) * torch.cos(t * x)
        else:
            v = torch.sin(t * x) * torch.cos(t * y)
        out.append((amplitude * v).to(dtype).contiguous())
    return out
