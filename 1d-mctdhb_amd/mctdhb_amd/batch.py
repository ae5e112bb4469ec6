"""Many independent MCTDHB solves on one GPU at once.

Each instance owns its configuration space, SIL workspace, orbital workspace
and state (orbitals, coefficients, h, g, rho1, rho2, rho1^-1) in HBM, and its
whole time step (reference driver.c:159-202) captured once as a HIP graph.
A batch replays the instances' graphs round-robin over a few streams, so the
small kernels of different instances run side by side on the compute units;
instances share no buffers, so each result is the same alone and in the mix.

`draw_ground_states` draws the uneven ground-state mix: N uniform in
[4, 64], M uniform in {2..6} (rejecting NC above a cap), Ng in {256, 512,
1024} grid intervals (periodic, grid_size Ng + 1), a double-well trap
V = a (x^2 - b^2)^2 with a ~ U[0.5, 2], b ~ U[1, 3], g log-uniform in
[0.01, 10], Hermite-Gauss orbitals plus 1e-3 noise and complex-normal
coefficients, imaginary time, dt = 0.005, Lanczos dimension 5.

The box is [-L, L) with V(L) = 200, not [-10, 10): the split step
integrates V inside its Runge-Kutta sub-step (orbital_integration.c:243-282),
which is stable only for V dt below about 2.8, and V(10) dt is 20 or more
for every draw, so on [-10, 10) the edge points blow up (the oracle's own
split step gives E/N in the thousands after 60 steps).  With V(L) dt = 1
the wells and their decaying tails lie well inside the box.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import inputs


@dataclass
class Draw:
    npar: int
    norb: int
    ngrid: int
    a: float
    b: float
    g: float
    seed: int

    @property
    def dim(self) -> int:
        return math.comb(self.npar + self.norb - 1, self.norb - 1)

    @property
    def half_width(self) -> float:
        """L with V(L) = a (L^2 - b^2)^2 = 200, at most 10."""
        return min(10.0, math.sqrt(self.b ** 2 + math.sqrt(200.0 / self.a)))


def draw_ground_states(count: int, seed: int = 11, max_dim: int = 200_000) -> list[Draw]:
    r = inputs.rng(seed, 4)
    out = []
    while len(out) < count:
        n = int(r.integers(4, 65))
        m = int(r.integers(2, 7))
        ng = int(r.choice([256, 512, 1024]))
        a = float(r.uniform(0.5, 2.0))
        b = float(r.uniform(1.0, 3.0))
        g = float(10 ** r.uniform(-2.0, 1.0))
        if math.comb(n + m - 1, m - 1) > max_dim:
            continue
        out.append(Draw(n, m, ng, a, b, g, seed * 100003 + len(out)))
    return out


# Above this many configurations an instance's initial coefficients are drawn
# on the device (inputs.coefficients_device), below it on the host.
DEVICE_DRAW_DIM = 1 << 16


def footprint_bytes(dim: int, norb: int, grid: int, lanczos: int = 5) -> int:
    """Device memory of one instance, upper estimate: coefficients and the
    Krylov vectors (complex128), the H.C output, occupation rows, the
    density partial sums, orbitals with their stage and FFT buffers, and a
    fixed allowance for the small tables and matrices."""
    return dim * (16 * (lanczos + 3) + 8) + (dim // 4096 + 2) * 16 * norb ** 4 + 16 * norb * grid * 24 + (1 << 20)


class Instance:
    """One imaginary-time solve, state resident on the GPU, step captured
    as a graph.  Spaces up to the one-workgroup SIL size run the FP64 step
    (its one-workgroup form is the faster one there); larger ones the mixed
    form (complex64 H.C, FP64 Lanczos scalars)."""

    def __init__(self, d: Draw, dt: float = 0.005, lanczos: int = 5, small_fp64_dim: int = 8192,
                 quench_g: float | None = None, fp64: bool = False):
        """quench_g: also capture a real-time step with interaction
        quench_g on the same state (`step_real`), for an interaction quench
        after the imaginary-time relaxation.  fp64: the FP64 form whatever
        the size (the twin a mixed result is checked against)."""
        import torch
        import mctdhb_amd as amd

        self.draw = d
        M, N = d.norb, d.npar
        grid = d.ngrid + 1
        L = d.half_width
        x = np.linspace(-L, L, grid)
        pot = d.a * (x ** 2 - d.b ** 2) ** 2
        self.prec = amd.FP64 if (fp64 or d.dim <= small_fp64_dim) else amd.MIXED
        self.space = amd.DeviceSpace(N, M, tables=False)
        self.sil = amd.DeviceSil(self.space, lanczos, self.prec, amd.INDEX_FLY)
        self.orbital = amd.DeviceOrbital(M, grid, x[0], x[-1], dt, d.g, pot, imagtime=True)
        orb = inputs.hermite_gauss(M, x, seed=d.seed, noise=1e-3)
        orb[:, -1] = orb[:, 0]
        dev = "cuda"
        # large synthetic states are drawn on the device (a host draw and copy
        # dominated the admission of a request with millions of configurations)
        coef = (inputs.coefficients_device(d.dim, d.seed, dev) if d.dim > DEVICE_DRAW_DIM else
                torch.from_numpy(inputs.coefficients(d.dim, d.seed)).to(dev))
        self.t = {"orb": torch.from_numpy(np.ascontiguousarray(orb.ravel())).to(dev), "coef": coef}
        for k, n in (("hob", M * M), ("hint", M ** 4), ("r1", M * M), ("r2", M ** 4), ("rinv", M * M)):
            self.t[k] = torch.zeros(n, dtype=torch.complex128, device=dev)
        t = self.t
        self.orbital.orthonormalize(t["orb"])
        self.orbital.orbital_matrices(t["orb"], t["hob"], t["hint"])
        amd.density_sync(self.space, t["coef"], t["r1"], t["r2"], t["rinv"])
        # the initial state is formed on the current stream: wait for that
        # stream alone, not for other solves' steps on their own streams
        torch.cuda.current_stream().synchronize()
        self.energy0 = self.energy_per_particle()  # E/N of the initial state
        self.graph = amd.StepGraph(self.space, self.sil, self.orbital, t["orb"], t["coef"], t["hob"], t["hint"],
                                   t["r1"], t["r2"], t["rinv"])
        self.orbital_real = self.graph_real = None
        self.quench_g = quench_g
        if quench_g is not None:
            self.orbital_real = amd.DeviceOrbital(M, grid, x[0], x[-1], dt, quench_g, pot, imagtime=False)
            self.graph_real = amd.StepGraph(self.space, self.sil, self.orbital_real, t["orb"], t["coef"], t["hob"],
                                            t["hint"], t["r1"], t["r2"], t["rinv"])

    def step(self, stream=None):
        self.graph.launch(stream)

    def step_real(self, stream=None):
        self.graph_real.launch(stream)

    def footprint_bytes(self) -> int:
        return footprint_bytes(self.draw.dim, self.draw.norb, self.draw.ngrid + 1)

    def _split(self, flat: np.ndarray):
        M = self.draw.norb
        m2, m4 = M * M, M ** 4
        return (flat[:m2].reshape(M, M), flat[m2:2 * m2].reshape(M, M),
                flat[2 * m2:2 * m2 + m4].reshape(M, M, M, M), flat[2 * m2 + m4:].reshape(M, M, M, M))

    def small_matrices(self) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        """rho1, h (M x M), rho2, g (M, M, M, M; flat index k + l M + s M^2
        + q M^3 on numpy axes (q, s, l, k)) on the host, in one
        device-to-host copy; call after a synchronisation."""
        import torch

        t = self.t
        return self._split(torch.cat([t["r1"], t["hob"], t["r2"], t["hint"]]).cpu().numpy())

    def snapshot(self, stream=None, tag: str = "imag"):
        """Enqueue on `stream` a copy of rho1, h, rho2 and g into pinned host
        memory, without waiting: read by health() once the stream has passed
        it (after the request's completion event).  tag: "imag" (during the
        relaxation), "relaxed" (before the first real-time step) or
        "quenched" (after it)."""
        import torch

        t = self.t
        with torch.cuda.stream(stream):
            src = torch.cat([t["r1"], t["hob"], t["r2"], t["hint"]])
            buf = torch.empty(src.shape, dtype=src.dtype, pin_memory=True)
            buf.copy_(src, non_blocking=True)
        self.snaps = getattr(self, "snaps", []) + [(tag, buf)]

    def energy_per_particle(self, mats=None) -> float:
        """E / N from rho1, rho2, h and g (reference total_energy,
        integrator_monitor.c:56-88); mats: small_matrices() already read.
        Call after a synchronisation."""
        r1, h, r2, g = mats if mats is not None else self.small_matrices()
        # rho2[k,l,s,q] g[k,l,q,s]: swap the two leading axes of g
        e = np.sum(r1 * h) + 0.5 * np.sum(r2 * g.transpose(1, 0, 2, 3))
        return float(e.real) / self.draw.npar

    # A real-time run whose energy moved by more than this fraction of the
    # energy right after the quench has not resolved the dynamics at its dt
    # (stable quenches of the served mix drift by 0.05-2 %).
    REAL_TIME_DRIFT = 0.1
    # A relaxation whose energy rose by more than this fraction between two
    # snapshots is unstable.  The split step's imaginary-time energy is not
    # strictly monotone (with weakly occupied orbitals it rises by 0.1 % and
    # more over 8 steps at dt = 0.005), while an unstable one rises by tens
    # of percent (g (N - 1) = 1185: 105 -> 120 -> 50 -> 97 over 10-step spans).
    RELAX_RISE = 0.1

    def health(self, tol: float = 1e-6) -> str:
        """Empty when the state is usable, else the first problem found:
        rho1 not finite, an energy above the grid's bound, an imaginary-time
        relaxation whose energy rose (by more than RELAX_RISE, relative) from
        one snapshot to the next (imaginary time lowers the energy: a rise means
        the split step is unstable at this dt, as when g (N - 1) dt is
        large, and the result is one of many trajectories rounding picks
        between), a quench whose real-time energy was not conserved to
        REAL_TIME_DRIFT, rho1's trace not N, or coefficients not of unit norm
        (every step keeps both, in imaginary time by renormalising and in
        real time by unitarity).  The energies come from the initial state,
        the snapshots (serve takes one every 32 relaxation steps and two
        around a quench's first real-time step) and, for a relaxation alone,
        the final state.  Call after a synchronisation."""
        import torch

        mats = self.small_matrices()
        r1 = mats[0]
        if not np.all(np.isfinite(r1)):
            return "rho1 not finite"
        e = self.energy_per_particle(mats)
        if abs(e) > self.energy_bound():
            return "energy above the grid's bound (diverged)"
        e0 = getattr(self, "energy0", None)
        snaps = getattr(self, "snaps", [])
        energy = lambda buf: self.energy_per_particle(self._split(buf.numpy()))
        relax = ([e0] if e0 is not None else []) + [energy(b) for tag, b in snaps if tag in ("imag", "relaxed")]
        if self.quench_g is None:
            relax.append(e)
        for a, b in zip(relax, relax[1:]):
            if b > a + self.RELAX_RISE * abs(a):
                return "imaginary-time energy rose (unstable step)"
        quenched = [b for tag, b in snaps if tag == "quenched"]
        if self.quench_g is not None and quenched:
            eq = energy(quenched[0])
            if not abs(e - eq) <= self.REAL_TIME_DRIFT * abs(eq):
                return "real-time energy not conserved (unstable step)"
        if abs(np.trace(r1).real - self.draw.npar) > tol * self.draw.npar:
            return "trace of rho1 is not N"
        nrm = float(torch.linalg.vector_norm(self.t["coef"]))  # on the device: one scalar comes back
        if not abs(nrm - 1.0) <= tol:
            return "coefficient norm is not 1"
        return ""

    def energy_bound(self) -> float:
        """An upper bound of |E/N| for any normalised state on this grid:
        max V, plus the largest grid kinetic energy (pi/dx)^2/2, plus
        |g|(N - 1) times the largest |phi|^2 of a normalised grid function
        (1/dx).  A result above it has diverged."""
        d = self.draw
        L = d.half_width
        dx = 2 * L / d.ngrid
        vmax = d.a * (L * L - d.b * d.b) ** 2
        g = max(abs(d.g), abs(self.quench_g)) if self.quench_g is not None else abs(d.g)
        return vmax + 0.5 * (np.pi / dx) ** 2 + g * (d.npar - 1) / dx

    def observables(self) -> tuple[float, np.ndarray, np.ndarray]:
        """(E/N, rho1, natural occupations in ascending order); call after a
        synchronisation."""
        mats = self.small_matrices()
        r1 = mats[0]
        return self.energy_per_particle(mats), r1, np.sort(np.linalg.eigvalsh(r1)) / self.draw.npar

    def occupations(self) -> np.ndarray:
        M = self.draw.norb
        r1 = self.t["r1"].cpu().numpy().reshape(M, M)
        return np.sort(np.linalg.eigvalsh(r1))[::-1] / self.draw.npar

    def close(self):
        """Frees the graphs first, then the workspaces they refer to."""
        self.graph.close()
        if self.graph_real is not None:
            self.graph_real.close()
            self.orbital_real.close()
        self.orbital.close()
        self.sil.close()
        self.space.close()


def errors_vs_fp64(got: tuple, ref: tuple) -> dict:
    """Errors of a result's observables against its FP64 twin's (both from
    Instance.observables): energy per particle (relative), natural
    occupations (largest relative deviation, and largest absolute one of
    the fractions n_k / N), rho1 (relative, Frobenius)."""
    (e, r1, occ), (e0, r10, occ0) = got, ref
    if not (np.isfinite(e) and np.isfinite(e0)):
        return {"energy_rel": None, "natural_occupations_rel": None, "natural_occupations_abs": None,
                "rho1_rel_fro": None}
    return {"energy_rel": abs(e - e0) / max(abs(e0), 1e-300),
            "natural_occupations_rel": float(np.max(np.abs(occ - occ0) / np.maximum(np.abs(occ0), 1e-300))),
            "natural_occupations_abs": float(np.max(np.abs(occ - occ0))),
            "rho1_rel_fro": float(np.linalg.norm(r1 - r10) / max(np.linalg.norm(r10), 1e-300))}


class Batch:
    """Instances stepped together: step k of every instance is enqueued on
    stream (instance index mod len(streams)) before step k + 1 of any.

    packed: the instances whose step qualifies (amd.pack_eligible: the
    one-workgroup FP64 step of the small spaces) run as an amd.Pack, one
    kernel launch per stage for each group of alike instances, on streams of
    their own; the others replay their step graphs as before."""

    def __init__(self, instances: list[Instance], nstreams: int = 4, packed: bool = False,
                 pack_streams: int = 8):
        import torch
        import mctdhb_amd as amd

        self.instances = instances
        self.streams = [torch.cuda.Stream() for _ in range(max(1, nstreams))]
        self.pack = None
        self.pack_streams = []
        self.graphs = [inst.graph for inst in instances]
        if packed:
            inpack = [i for i in instances if amd.pack_eligible(i.space, i.sil, i.orbital)]
            if inpack:
                t = lambda i: (i.space, i.sil, i.orbital, i.t["orb"], i.t["coef"], i.t["hob"], i.t["hint"],
                               i.t["r1"], i.t["r2"], i.t["rinv"])
                self.pack = amd.Pack([t(i) for i in inpack])
                self.pack_streams = [torch.cuda.Stream() for _ in range(max(1, min(pack_streams,
                                                                                    self.pack.groups)))]
                ids = {id(i) for i in inpack}
                self.graphs = [i.graph for i in instances if id(i) not in ids]
        self.packed_count = len(instances) - len(self.graphs)

    def run(self, steps: int):
        """steps rounds: the pack's groups (one graph of stage launches per
        group and round) and the other instances' step graphs, instance i's
        on stream i mod len(streams); one library call each."""
        import mctdhb_amd as amd

        if self.pack is not None:
            self.pack.step(self.pack_streams, steps)
        amd.launch_graphs(self.graphs, self.streams, steps)

    def synchronize(self):
        for s in self.streams + self.pack_streams:
            s.synchronize()

    def close(self):
        if self.pack is not None:
            self.pack.close()
            self.pack = None
