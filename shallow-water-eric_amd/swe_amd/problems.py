"""Problem set-ups for the benchmark workloads and the parity cases.

* :func:`paraboloid`   planar free surface oscillating in a paraboloid: the reference's test 5
  (init: SOURCES/condlim.f90:151-161, sol_anal: 429-456), first-order graph viscosity.
* :func:`dam_break`    dam break on a flat [0,100]x[0,20] rectangle with a jittered structured
  mesh, Manning friction (the reference's friction source, dispatched for type_test 10,
  SOURCES/update.f90:614-616), FCT + alpha limiter, reflecting walls.
* :func:`bumps`        still water over seeded Gaussian bumps on a jittered-lattice Delaunay mesh,
  with a sinusoidal free-surface inflow on the west side; FCT + alpha limiter.

Node numbering is the reference loader's (first-touch renumbering of the generated element list),
so a problem built here and the reference fed with the same ``.FEM`` file see the same arrays.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import mesh as M
from .solver import BC_INFLOW, BC_NONE, BC_WALLS, Solver, assemble


@dataclasses.dataclass
class Problem:
    name: str
    mesh: M.Mesh                 # renumbered as the reference loader does
    ia: np.ndarray
    ja: np.ndarray
    diag: np.ndarray
    lumped: np.ndarray
    cij: np.ndarray
    bath: np.ndarray
    u0: np.ndarray               # (np, 3)
    params: dict
    h_sides: tuple = ()
    ux_sides: tuple = ()
    uy_sides: tuple = ()
    bc: tuple = (BC_NONE, 0.0, 1.0, 0.0)   # (kind, amp, period, level)
    nsteps: int = 1
    reference_init: bool = False           # True when the reference's own init covers the case
    tfinal: float | None = None            # final time of a served request

    @property
    def np(self):
        return self.mesh.np

    @property
    def nnz(self):
        return int(self.ja.size)

    def dirichlet(self):
        return (M.dirichlet_nodes(self.mesh, self.h_sides), M.dirichlet_nodes(self.mesh, self.ux_sides),
                M.dirichlet_nodes(self.mesh, self.uy_sides))

    def solver(self, device_setup=False, **overrides) -> Solver:
        """A solver for this case; device_setup=True assembles and lays the mesh out on the GPU
        (Solver.from_elements) instead of taking the host assembly."""
        prm = dict(self.params)
        prm.update(overrides)
        hd, uxd, uyd = self.dirichlet()
        if device_setup:
            s = Solver.from_elements(self.mesh.jj, self.mesh.rr, self.bath, h_dir=hd, ux_dir=uxd, uy_dir=uyd, **prm)
        else:
            s = Solver(self.ia, self.ja, self.cij, self.lumped, self.bath, h_dir=hd, ux_dir=uxd, uy_dir=uyd, **prm)
        s.set_bc(*self.bc)
        return s


def _finish(name, raw_mesh, bath_fn, u0_fn, params, host_assembly=True, **kw):
    """host_assembly=False leaves the matrices (ia, ja, diag, lumped, cij) out: such a problem is
    solved through solver(device_setup=True), which assembles on the GPU."""
    m = M.renumber_first_touch(raw_mesh)
    if host_assembly:
        ia, ja, diag, _mass, lumped, cij = assemble(m.jj, m.rr)
    else:
        ia = ja = diag = lumped = cij = None
    bath = bath_fn(m.rr)
    u0 = u0_fn(m.rr, bath)
    return Problem(name=name, mesh=m, ia=ia, ja=ja, diag=diag, lumped=lumped, cij=cij, bath=bath,
                   u0=np.ascontiguousarray(u0), params=params, **kw)


# ----------------------------------------------------------------------------------------------
def paraboloid_state(rr, t, htiny):
    """sol_anal of test 5 (SOURCES/condlim.f90:429-456), evaluated in the reference's order."""
    g, a, h0, hL, eta = 9.81, 1.0, 0.1, 4.0, 0.5
    omega = math.sqrt(2 * g * h0) / a
    x, y = rr[:, 0], rr[:, 1]
    bathi = -h0 * (1.0 - ((x - hL / 2.0) ** 2 + (y - hL / 2.0) ** 2) / a ** 2)
    c, s = math.cos(omega * t), math.sin(omega * t)
    base = np.maximum((eta * h0 / a ** 2) * (2.0 * (x - hL / 2.0) * c + 2.0 * (y - hL / 2.0) * s) - bathi, 0.0)
    u = np.empty((rr.shape[0], 3))
    u[:, 0] = base + htiny
    u[:, 1] = -eta * omega * s * base
    u[:, 2] = eta * omega * c * base
    return u


def paraboloid(n=81, cfl=0.25, viscosity="viscous", alpha=False, viscous_type=2) -> Problem:
    """Test 5 on an n-by-n P1 grid of [0,4]^2 split along (x,y)->(x+h,y+h): the first-order
    graph-viscosity configuration (81x81 = 6,561 nodes, CFL 0.25, 2000 steps)."""
    g, h0, a, hL = 9.81, 0.1, 1.0, 4.0
    max_water_h = hL
    htiny = 1.0e-4 * max_water_h

    def bath(rr):
        return -h0 * (1.0 - ((rr[:, 0] - hL / 2.0) ** 2 + (rr[:, 1] - hL / 2.0) ** 2) / a ** 2)

    params = dict(gravity=g, htiny=htiny, cfl=cfl, max_water_h=max_water_h, viscosity=viscosity,
                  viscous_type=viscous_type, alpha_limit=alpha, type_test=5, mannings=0.0)
    return _finish("paraboloid", M.structured_mesh(n, n, 0.0, 4.0, 0.0, 4.0), bath,
                   lambda rr, z: paraboloid_state(rr, 0.0, htiny), params, nsteps=2000, reference_init=True)


# ----------------------------------------------------------------------------------------------
def dam_break(nx=2241, ny=449, jitter=0.2, seed=20261202, cfl=0.25, mannings=0.02,
              viscosity="fct", alpha=True) -> Problem:
    """Dam break on [0,100]x[0,20]: h=1 for x<50, dry elsewhere, q=0, flat bottom, Manning n=0.02,
    walls (qx=0 on x=0 and x=100, qy=0 on y=0 and y=20).  2241x449 = 1,006,209 nodes, 500 steps."""
    params = dict(gravity=9.81, htiny=1.0e-4 * 1.0, cfl=cfl, max_water_h=1.0, viscosity=viscosity,
                  viscous_type=2, alpha_limit=alpha, type_test=10, mannings=mannings)

    def u0(rr, z):
        u = np.zeros((rr.shape[0], 3))
        u[:, 0] = np.where(rr[:, 0] < 50.0, 1.0, 0.0)
        return u
    return _finish("dam_break", M.structured_mesh(nx, ny, 0.0, 100.0, 0.0, 20.0, jitter=jitter, seed=seed),
                   lambda rr: np.zeros(rr.shape[0]), u0, params, ux_sides=(2, 4), uy_sides=(1, 3),
                   bc=(BC_WALLS, 0.0, 1.0, 0.0), nsteps=500)


# ----------------------------------------------------------------------------------------------
def sgn_solitary_wave(rr, g=9.81):
    """The SGN model's initial state of test 13: the Favrie-Gavrilyuk solitary wave of
    sol_anal(1..5, rr, t=0) (SOURCES/condlim.f90:751-827): h1 = 0.10, h2 = 0.11, crest at x0 = 2.75,
    celerity D = sqrt(g h2); components h, h*u, h*v, h*eta = h^2, h*w."""
    h1, h2, x0 = 10.0 / 100.0, 11.0 / 100.0, 2.75
    D = math.sqrt(g * h2)
    zz = math.sqrt((3.0 * (h2 - h1)) / (h2 * h1 ** 2.0))
    arg = 1.0 / 2.0 * zz * (rr[:, 0] - x0)
    sech2 = (1.0 / np.cosh(arg)) ** 2.0
    h = np.maximum(h1 + (h2 - h1) * sech2, 0.0)
    u = np.empty((rr.shape[0], 5))
    u[:, 0] = h
    u[:, 1] = h * D - h1 * D
    u[:, 2] = h * 0.0
    u[:, 3] = h * h
    u[:, 4] = -D * h1 * ((h1 - h2) * zz * sech2 * np.tanh(arg))
    return u


def sgn_canal(nx=6401, ny=641, jitter=0.2, seed=20261202, viscosity="viscous", alpha=True, cfl=0.08,
              lambda_sgn=9.0) -> Problem:
    """The SGN model (test 13) at scale: the reference's canal [0,10]x[0,1]
    (MESHES/SGN_CANAL/rectangle.*.FEM, data lambdaSGN = 9, CFL 0.08, max_water_h = 0.11) on an
    nx-by-ny jittered lattice (6401x641 = 4,103,041 nodes), the solitary wave of sgn_solitary_wave,
    flat bottom, no Dirichlet nodes (the wave stays far from the ends over a bench run), first-order
    viscous scheme with the alpha limiter."""
    g, max_water_h = 9.81, 0.11
    params = dict(gravity=g, htiny=1.0e-4 * max_water_h, cfl=cfl, max_water_h=max_water_h, viscosity=viscosity,
                  viscous_type=2, alpha_limit=alpha, type_test=13, mannings=0.0, lambda_sgn=lambda_sgn)
    return _finish("sgn_canal", M.structured_mesh(nx, ny, 0.0, 10.0, 0.0, 1.0, jitter=jitter, seed=seed),
                   lambda rr: np.zeros(rr.shape[0]), lambda rr, z: sgn_solitary_wave(rr, g), params, nsteps=200,
                   reference_init=True)


# ----------------------------------------------------------------------------------------------
def bump_field(rr, L, nbumps=64, seed=20261202):
    """z = sum_k A_k exp(-|x-c_k|^2/(2 w_k^2)), A ~ U[0,0.8], w ~ U[0.02L, 0.1L], c ~ U[0,L]^2."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.0, 0.8, nbumps)
    w = rng.uniform(0.02 * L, 0.1 * L, nbumps)
    c = rng.uniform(0.0, L, (nbumps, 2))
    z = np.zeros(rr.shape[0])
    for k in range(nbumps):
        d2 = (rr[:, 0] - c[k, 0]) ** 2 + (rr[:, 1] - c[k, 1]) ** 2
        z += A[k] * np.exp(-d2 / (2.0 * w[k] ** 2))
    return z


def bumps(n_side=4097, L=1000.0, nbumps=64, seed=20261202, cfl=0.25, amp=0.05, period=20.0,
          viscosity="fct", alpha=True) -> Problem:
    """Still water of level 1 over 64 Gaussian bumps (dry islands where z > 1) on a cell-split
    jittered lattice of n_side^2 points of [0,L]^2 (each lattice cell split along the diagonal the
    in-circle test picks: the Delaunay triangulation of the jittered points, with the lattice's
    structured topology and uniform valence); free-surface inflow
    h = max(1 - z + amp*sin(2*pi*t/period), 0) on the west side, walls elsewhere.  n_side = 4097
    gives 16,785,409 nodes (the lattice form of the 16.8M-node workload; bumps_unstructured is the
    unstructured one), 200 steps."""
    params = dict(gravity=9.81, htiny=1.0e-4 * 1.0, cfl=cfl, max_water_h=1.0, viscosity=viscosity,
                  viscous_type=2, alpha_limit=alpha, type_test=4, mannings=0.0)

    def u0(rr, z):
        u = np.zeros((rr.shape[0], 3))
        u[:, 0] = np.maximum(1.0 - z, 0.0)
        return u
    return _finish("bumps", M.jittered_delaunay_mesh(n_side, L, jitter=0.25, seed=seed),
                   lambda rr: bump_field(rr, L, nbumps, seed), u0, params,
                   h_sides=(4,), ux_sides=(2,), uy_sides=(1, 3), bc=(BC_INFLOW, amp, period, 1.0), nsteps=200)


def bumps_unstructured(n_nodes=16_785_409, L=1000.0, nbumps=64, seed=20261202, cfl=0.25, amp=0.05, period=20.0,
                       viscosity="fct", alpha=True, host_assembly=True) -> Problem:
    """The 16.8M-node workload on an unstructured mesh: still water of level 1 over 64 Gaussian
    bumps (dry islands where z > 1), the sinusoidal west inflow, walls elsewhere, FCT + alpha,
    200 steps, on the Delaunay triangulation of a seeded Poisson-disk point set of about n_nodes
    nodes (mesh.poisson_delaunay_mesh), renumbered by the reference loader's first-touch rule."""
    params = dict(gravity=9.81, htiny=1.0e-4 * 1.0, cfl=cfl, max_water_h=1.0, viscosity=viscosity,
                  viscous_type=2, alpha_limit=alpha, type_test=4, mannings=0.0)

    def u0(rr, z):
        u = np.zeros((rr.shape[0], 3))
        u[:, 0] = np.maximum(1.0 - z, 0.0)
        return u
    return _finish("bumps_pd", M.poisson_delaunay_mesh(L, n_nodes=n_nodes, seed=seed),
                   lambda rr: bump_field(rr, L, nbumps, seed), u0, params,
                   h_sides=(4,), ux_sides=(2,), uy_sides=(1, 3), bc=(BC_INFLOW, amp, period, 1.0), nsteps=200,
                   host_assembly=host_assembly)


# ----------------------------------------------------------------------------------------------
def family_member(n_nodes, seed, kind="bumps", L=1000.0, tfinal=None, amp=0.05, period=20.0, zmax=0.7,
                  cfl=0.25, mesh="poisson") -> Problem:
    """One mesh of the 16.8M-node workload's family at another size: an unstructured mesh of about
    n_nodes nodes on [0,L]^2 (the Delaunay triangulation of a Poisson-disk sample, seeded; mesh =
    "lattice" gives the cell-split jittered lattice instead) with its own 64 Gaussian bumps, scaled down
    where needed so that max z <= zmax: the bumps stay submerged (the reference scheme stops with
    'Min h<0' within a few hundred steps on many coarse meshes with emerging islands, so the long
    served runs keep the islands out).  zmax = None keeps the bumps as drawn: where they sum above the
    still-water level they emerge as dry islands (h = 0 there from the start).  Kinds:
      'bumps'   still water of level 1 with the sinusoidal west inflow, walls elsewhere;
      'dam'     dam break over the bumps: level 1 on x < L/2, level 0.8 beyond, q = 0, walls;
      'drydam'  the dry-bed dam break of the 1M-node workload on this mesh: flat bed, h = 1 on
                x < L/2 and 0 beyond, Manning n = 0.02 (type_test 10), walls.
    FCT + alpha limiter; type_test 4 (no source term) except for 'drydam'."""
    n_side = max(8, int(round(math.sqrt(n_nodes))))
    params = dict(gravity=9.81, htiny=1.0e-4 * 1.0, cfl=cfl, max_water_h=1.0, viscosity="fct",
                  viscous_type=2, alpha_limit=True, type_test=4, mannings=0.0)

    def bath(rr):
        if kind == "drydam":
            return np.zeros(rr.shape[0])
        z = bump_field(rr, L, 64, seed)
        return z * (zmax / z.max()) if zmax is not None and z.max() > zmax else z

    walls = dict(ux_sides=(2, 4), uy_sides=(1, 3), bc=(BC_WALLS, 0.0, 1.0, 0.0))
    if kind == "bumps":
        def u0(rr, z):
            u = np.zeros((rr.shape[0], 3))
            u[:, 0] = np.maximum(1.0 - z, 0.0)
            return u
        bc = dict(h_sides=(4,), ux_sides=(2,), uy_sides=(1, 3), bc=(BC_INFLOW, amp, period, 1.0))
    elif kind == "basin":
        cx, cy = np.random.default_rng(seed + 7).uniform(0.2 * L, 0.8 * L, 2)

        def u0(rr, z):
            u = np.zeros((rr.shape[0], 3))
            eta = 1.0 + 0.05 * np.exp(-((rr[:, 0] - cx) ** 2 + (rr[:, 1] - cy) ** 2) / (2.0 * (0.05 * L) ** 2))
            u[:, 0] = np.maximum(eta - z, 0.0)
            return u
        bc = walls
    elif kind == "dam":
        def u0(rr, z):
            u = np.zeros((rr.shape[0], 3))
            u[:, 0] = np.where(rr[:, 0] < 0.5 * L, np.maximum(1.0 - z, 0.0), np.maximum(0.8 - z, 0.0))
            return u
        bc = walls
    elif kind == "drydam":
        params.update(type_test=10, mannings=0.02)

        def u0(rr, z):
            u = np.zeros((rr.shape[0], 3))
            u[:, 0] = np.where(rr[:, 0] < 0.5 * L, 1.0, 0.0)
            return u
        bc = walls
    else:
        raise ValueError("unknown kind %r" % kind)
    if mesh == "poisson":
        raw = M.poisson_delaunay_mesh(L, n_nodes=n_nodes, seed=seed)
    else:
        raw = M.jittered_delaunay_mesh(n_side, L, jitter=0.25, seed=seed)
    return _finish("%s_%d" % (kind, raw.np), raw, bath, u0, params, nsteps=1000, tfinal=tfinal, **bc)


def batch_family_specs(count=256, nmin=5000, nmax=500000, seed=20261202, dam_fraction=0.5, island_fraction=0.25):
    """Sizes (log-uniform in [nmin, nmax]), seeds and kinds of the batched workload's instances:
    dam breaks ('dam') with probability dam_fraction, still water ('basin') otherwise; with
    probability island_fraction an instance keeps its bumps as drawn (zmax None: dry islands where
    they emerge), the others have them scaled under the still-water level."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(nmin), math.log(nmax), count))
    kinds = np.where(rng.uniform(size=count) < dam_fraction, "dam", "basin")
    seeds = rng.integers(1, 2 ** 31 - 1, count)
    # drawn after the others so that sizes, kinds and seeds stay those of the same seed without islands
    islands = rng.uniform(size=count) < island_fraction
    return [dict(n_nodes=int(n), seed=int(s), kind=str(k), zmax=None if isl else 0.7)
            for n, s, k, isl in zip(sizes, seeds, kinds, islands)]


def serve_request_specs(count, rate, seed=20261202, median=100000, sigma=1.0, nmax=2000000, nmin=2000,
                        dry_fraction=0.2, period=20.0, L=1000.0):
    """A Poisson arrival stream: arrival times (s), sizes lognormal (median, sigma, clipped to
    [nmin, nmax]), final times U[1,20] wave periods, dry-bed dam breaks ('drydam') with probability
    dry_fraction, the rest a wave in a still-water basin over their own bumps ('basin')."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    sizes = np.clip(np.exp(math.log(median) + sigma * rng.standard_normal(count)), nmin, nmax)
    tfinal = rng.uniform(1.0, 20.0, count) * period
    kinds = np.where(rng.uniform(size=count) < dry_fraction, "drydam", "basin")
    seeds = rng.integers(1, 2 ** 31 - 1, count)
    return [dict(arrival=float(a), n_nodes=int(n), tfinal=float(t), kind=str(k), seed=int(s), L=L, period=period)
            for a, n, t, k, s in zip(arrivals, sizes, tfinal, kinds, seeds)]


def build_family(specs, workers=1):
    """Problems of a list of specs (family_member keyword sets), in parallel processes if workers > 1."""
    keys = ("n_nodes", "seed", "kind", "L", "tfinal", "period", "mesh", "zmax")
    args = [{k: s[k] for k in keys if k in s} for s in specs]
    if workers <= 1 or len(args) < 2:
        return [family_member(**a) for a in args]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_family_member_kw, args, chunksize=1)


def _family_member_kw(kw):
    M.MESHGEN_THREADS = 1       # one generator thread per worker process
    return family_member(**kw)
