"""Problem set-up on the host: structured quad meshes, boundary node types,
initial conditions and configuration, in the dataAPI layout.

This replaces the reference's deck reader / mesh generator / IC set-up
(src/packages/setup, out of the hot path) for the workloads the hot path is
measured on.  Conventions follow the reference:

* node types (setup_mesh.F90:1379-1584): interior nodes carry their region
  number (> 0); boundary nodes carry -bc with SLIPX=1, SLIPY=2, WALL=3,
  FREE=8 (setup_mesh.F90:324-341); corners joining SLIPX and SLIPY become -3
  (setup_mesh.F90:886-1000);
* element corners are counter-clockwise, face f joins corners f and f+1
  (dataAPI_params.f90:27-28 IFCND);
* energy given per unit mass; Sedov's deposit uses energy_scale='mass'
  semantics e = E/(rho*V) (setup_IC.f90:621-625);
* defaults of *_cf_defaults (global_config.f90:37-40, hydro_config.f90:43-49,
  eos_config.f90:42-43, time_config.f90:37-42, ale_config.f90:42-49).
"""
from dataclasses import dataclass, field
import math

import numpy as np

from . import abi

BC = {"SLIPX": 1, "SLIPY": 2, "WALL": 3, "FREE": 8}

DEFAULT_CONFIG = dict(
    zcut=1.0e-8, zerocut=1.0e-40, dencut=1.0e-6, accut=1.0e-6,
    cvisc1=0.5, cvisc2=0.75, cfl_sf=0.5, div_sf=0.25, zhg=0, zsp=0, ztq=0, pad0=0,
    ccut=1.0e-6, pcut=1.0e-8,
    time_start=0.0, time_end=1.0, dt_g=1.02, dt_min=1.0e-8, dt_max=1.0e-1,
    dt_initial=1.0e-5,
    mintime=np.finfo(np.float64).max, maxtime=np.finfo(np.float64).max, sf=0.5,
    adv_type=1, zexist=0, zon=0, zeul=0)


@dataclass
class Problem:
    name: str
    ielnd: np.ndarray          # (4*nel) int32, 1-based, corners of element e at 4e..4e+3
    ndx: np.ndarray
    ndy: np.ndarray
    indtype: np.ndarray        # (nnd) int32
    ielreg: np.ndarray         # (nel) int32 1-based
    ielmat: np.ndarray         # (nel) int32 1-based
    density: np.ndarray
    energy: np.ndarray
    ndu: np.ndarray
    ndv: np.ndarray
    nmat: int
    nreg: int
    eos_itype: np.ndarray      # (nmat+1) int32, [0] = void
    eos_param: np.ndarray      # (6*(nmat+1)) float64, material m at [6m:6m+6]
    kappareg: np.ndarray = None
    pmeritreg: np.ndarray = None
    zdtnotreg: np.ndarray = None
    zmidlength: np.ndarray = None
    config: dict = field(default_factory=lambda: dict(DEFAULT_CONFIG))
    nsteps: int = 1
    stop_at_time_end: int = 0
    dims: tuple = None
    # mixed cells (ncp > 0): the flattened component lists of mix_dr_flatten
    # (mix_dr_list.f90), 1-based contents; ielmat(iel) = -imx for mixed cells
    imxel: np.ndarray = None       # (nmx) element of each mixed cell
    imxfcp: np.ndarray = None      # (nmx) first component
    imxncp: np.ndarray = None      # (nmx) number of components
    icpmat: np.ndarray = None      # (ncp) material of each component
    icpnext: np.ndarray = None     # (ncp) next component (0: last)
    icpprev: np.ndarray = None     # (ncp) previous component (-imx: first)
    frvolume: np.ndarray = None    # (ncp) volume fraction
    cpdensity: np.ndarray = None   # (ncp)
    cpenergy: np.ndarray = None    # (ncp)
    # ALE node relaxation (bl_set_relaxation): (type, nsweep, om, tol)
    relax: tuple = None

    def __post_init__(self):
        if self.kappareg is None:
            self.kappareg = np.zeros(self.nreg)
        if self.pmeritreg is None:
            self.pmeritreg = np.zeros(self.nreg)
        if self.zdtnotreg is None:
            self.zdtnotreg = np.zeros(self.nreg, np.int32)
        if self.zmidlength is None:
            self.zmidlength = np.zeros(self.nreg, np.int32)
        # hydro_cf_rationalise (hydro_config.f90:349,368): zhg/zsp from tables
        self.config["zhg"] = int(np.any(self.kappareg > 0.0))
        self.config["zsp"] = int(np.any(self.pmeritreg > 0.0))
        # ale_cf_rationalise (ale_config.f90:539-551): Eulerian implies ALE on
        if self.config.get("zeul"):
            self.config["zexist"] = 1
            self.config["mintime"] = -np.finfo(np.float64).max
            self.config["maxtime"] = np.finfo(np.float64).max

    @property
    def nel(self):
        return len(self.ielreg)

    @property
    def nnd(self):
        return len(self.ndx)

    @property
    def ncp(self):
        return 0 if self.icpmat is None else len(self.icpmat)

    @property
    def nmx(self):
        return 0 if self.imxel is None else len(self.imxel)

    def sizes(self):
        s = abi.Sizes()
        s.nel = s.nel1 = s.nel2 = self.nel
        s.nnd = s.nnd1 = s.nnd2 = self.nnd
        s.nsz = max(self.nel, self.nnd)
        s.nmat, s.nreg = self.nmat, self.nreg
        s.ncp, s.nmx = self.ncp, self.nmx
        # component capacities as mix_dr_addcp / mix_dr_addel first size them
        # (INCR = 5%, mix_dr_list.f90:53-57,110-114), and at least 4*ncp: the
        # reference registers cpviscx/cpviscy with one entry per component
        # (data.f90:360-361) but hydro_kn_viscinit and hydro_kn_getforceq
        # address four per component
        s.mcp = max(4 * self.ncp, int(0.05 * self.nmat * self.nel)) if self.ncp else 0
        s.mmx = max(self.nmx, int(0.05 * self.nel)) if self.nmx else 0
        return s

    def set_mixed(self, cells):
        """Make cells mixed.  cells: {element (0-based): [(material, frvolume,
        density, energy), ...]}.  The lists are laid out as mix_dr_flatten
        leaves them (mix_kn_list.f90:230-330): mixed cells in ascending element
        order, each cell's components contiguous and in ascending material
        order, icpnext/icpprev chaining them, ielmat(iel) = -imx."""
        imxel, imxfcp, imxncp, icpmat, frv, rho, en = [], [], [], [], [], [], []
        for e in sorted(cells):
            comps = sorted(cells[e], key=lambda c: c[0])
            imxel.append(e + 1)
            imxfcp.append(len(icpmat) + 1)
            imxncp.append(len(comps))
            for m, f, r, u in comps:
                icpmat.append(m); frv.append(f); rho.append(r); en.append(u)
            self.ielmat[e] = -len(imxel)
        ncp = len(icpmat)
        nxt = np.zeros(ncp, np.int32)
        prv = np.zeros(ncp, np.int32)
        for imx, (f, n) in enumerate(zip(imxfcp, imxncp), start=1):
            for k in range(n):
                icp = f + k
                nxt[icp - 1] = icp + 1 if k < n - 1 else 0
                prv[icp - 1] = icp - 1 if k > 0 else -imx
        self.imxel = np.array(imxel, np.int32)
        self.imxfcp = np.array(imxfcp, np.int32)
        self.imxncp = np.array(imxncp, np.int32)
        self.icpmat = np.array(icpmat, np.int32)
        self.icpnext, self.icpprev = nxt, prv
        self.frvolume = np.array(frv, np.float64)
        self.cpdensity = np.array(rho, np.float64)
        self.cpenergy = np.array(en, np.float64)

    def cconfig(self):
        c = abi.Config()
        for k, v in self.config.items():
            setattr(c, k, v)
        return c


def _corner_code(a, b):
    """Corner BC code from the two adjacent side codes (setup_mesh.F90:886-1000)."""
    s = {a, b}
    if s == {1, 2}:
        return 3
    if 8 in s and len(s) == 2:
        (o,) = s - {8}
        return o
    return max(a, b)


def structured_mesh(nx, ny, x0, x1, y0, y1, bcs=("SLIPY", "SLIPX", "SLIPY", "SLIPX"),
                    jitter=0.0, seed=12345):
    """nx*ny quads on [x0,x1]x[y0,y1]; bcs = (bottom, right, top, left).

    Node (i,j) has 1-based id 1+i+j*(nx+1); element (i,j) has id 1+i+j*nx with
    corners (i,j),(i+1,j),(i+1,j+1),(i,j+1).  Interior nodes are displaced by
    U(-jitter,jitter)*cell width in x and y (numpy default_rng(seed)).
    """
    xs = x0 + (x1 - x0) * np.arange(nx + 1, dtype=np.float64) / nx
    ys = y0 + (y1 - y0) * np.arange(ny + 1, dtype=np.float64) / ny
    X, Y = np.meshgrid(xs, ys)                   # (ny+1, nx+1)
    if jitter:
        rng = np.random.default_rng(seed)
        hx, hy = (x1 - x0) / nx, (y1 - y0) / ny
        jx = rng.uniform(-jitter, jitter, size=(ny - 1, nx - 1)) * hx
        jy = rng.uniform(-jitter, jitter, size=(ny - 1, nx - 1)) * hy
        X[1:-1, 1:-1] += jx
        Y[1:-1, 1:-1] += jy
    ndx = np.ascontiguousarray(X.ravel())
    ndy = np.ascontiguousarray(Y.ravel())
    i = np.arange(nx)
    j = np.arange(ny)
    I, J = np.meshgrid(i, j)
    n00 = (I + J * (nx + 1)).ravel()
    ielnd = np.stack([n00, n00 + 1, n00 + 1 + (nx + 1), n00 + (nx + 1)], axis=1)
    ielnd = (ielnd + 1).astype(np.int32).ravel()
    bottom, right, top, left = (BC[b] for b in bcs)
    code = np.zeros((ny + 1, nx + 1), np.int32)
    code[0, :] = bottom
    code[-1, :] = top
    code[:, 0] = left
    code[:, -1] = right
    code[0, 0] = _corner_code(bottom, left)
    code[0, -1] = _corner_code(bottom, right)
    code[-1, 0] = _corner_code(top, left)
    code[-1, -1] = _corner_code(top, right)
    indtype = np.where(code > 0, -code, 1).astype(np.int32).ravel()
    return ielnd, ndx, ndy, indtype


def element_volume(ielnd, ndx, ndy, e):
    """V of element e (0-based) with the exact operation order of
    geometry_kn_getiso/getvolume (geometry_kn_volume.f90:49-85)."""
    n = [int(ielnd[4 * e + k]) - 1 for k in range(4)]
    x = [float(ndx[k]) for k in n]
    y = [float(ndy[k]) for k in n]
    a1 = 0.25 * (-x[0] + x[1] + x[2] - x[3])
    a3 = 0.25 * (-x[0] - x[1] + x[2] + x[3])
    b1 = 0.25 * (-y[0] + y[1] + y[2] - y[3])
    b3 = 0.25 * (-y[0] - y[1] + y[2] + y[3])
    return 4.0 * (a1 * b3 - a3 * b1)


def centroids(ielnd, ndx, ndy):
    n = ielnd.reshape(-1, 4) - 1
    return ndx[n].mean(axis=1), ndy[n].mean(axis=1)


def ideal_eos(gammas):
    nmat = len(gammas)
    itype = np.zeros(nmat + 1, np.int32)
    param = np.zeros(abi.NEOSPARAM * (nmat + 1))
    for m, g in enumerate(gammas, start=1):
        itype[m] = 1
        param[abi.NEOSPARAM * m] = g
    return itype, param


def sod(nx=200, ny=20, length=1.0, height=0.1, time_end=0.2, nsteps=None, eulerian=False,
        **cfg):
    """Sod shock tube: left rho=1 p=1, right rho=0.125 p=0.1, gamma=1.4, u=0;
    interface at mid-length; SLIPY bottom/top, SLIPX ends.  Regions 1 (left)
    and 2 (right), one material."""
    ielnd, ndx, ndy, indtype = structured_mesh(nx, ny, 0.0, length, 0.0, height)
    xc, _ = centroids(ielnd, ndx, ndy)
    left = xc < 0.5 * length
    nel = nx * ny
    density = np.where(left, 1.0, 0.125)
    energy = np.where(left, 2.5, 2.0)   # e = p/((gamma-1) rho)
    itype, param = ideal_eos([1.4])
    config = dict(DEFAULT_CONFIG, time_end=time_end, **cfg)
    if eulerian:
        config["zeul"] = 1
    return Problem(
        name="sod_%dx%d%s" % (nx, ny, "_eul" if eulerian else ""),
        ielnd=ielnd, ndx=ndx, ndy=ndy, indtype=indtype,
        ielreg=np.where(left, 1, 2).astype(np.int32), ielmat=np.ones(nel, np.int32),
        density=density, energy=energy, ndu=np.zeros(len(ndx)), ndv=np.zeros(len(ndx)),
        nmat=1, nreg=2, eos_itype=itype, eos_param=param, config=config,
        nsteps=nsteps if nsteps is not None else 100000,
        stop_at_time_end=1 if nsteps is None else 0, dims=(nx, ny))


def sedov(n=1024, length=1.2, jitter=0.1, seed=12345, energy_total=0.25, gamma=1.4,
          nsteps=500, kappa=0.0, **cfg):
    """Sedov blast on [0,L]^2, rho=1, E deposited in the corner element at the
    origin as e = E/(rho*V1), e = 0 elsewhere; interior nodes jittered by
    U(-jitter,jitter) cell widths; reflective walls (SLIPY bottom/top, SLIPX
    left/right)."""
    ielnd, ndx, ndy, indtype = structured_mesh(n, n, 0.0, length, 0.0, length,
                                               jitter=jitter, seed=seed)
    nel = n * n
    density = np.ones(nel)
    energy = np.zeros(nel)
    v1 = element_volume(ielnd, ndx, ndy, 0)
    energy[0] = energy_total / (1.0 * v1)
    itype, param = ideal_eos([gamma])
    config = dict(DEFAULT_CONFIG, time_end=1.0e30, **cfg)
    return Problem(
        name="sedov_%dx%d" % (n, n), ielnd=ielnd, ndx=ndx, ndy=ndy, indtype=indtype,
        ielreg=np.ones(nel, np.int32), ielmat=np.ones(nel, np.int32), density=density,
        energy=energy, ndu=np.zeros(len(ndx)), ndv=np.zeros(len(ndx)), nmat=1, nreg=1,
        eos_itype=itype, eos_param=param, kappareg=np.array([kappa]), config=config,
        nsteps=nsteps, dims=(n, n))


# ---- ensembles ----------------------------------------------------------

ENSEMBLE_CONFIG = dict(DEFAULT_CONFIG, time_end=1.0e30, dt_initial=1.0e-3)


def ensemble_params(ninst=512, seed=2025, nmin=64, nmax=512):
    """Instances of a Sod/Sedov ensemble, one dict each: a square n x n mesh
    with n uniform over nmin..nmax, Sod or Sedov with even odds, gamma
    U[1.3, 1.67]; Sod: left rho = 1, p = 1 and right rho, p lower by ratios
    log-uniform over [2, 20] and [2, 100] on [0,1]^2; Sedov: rho = 1, E =
    0.25 in the corner element on [0,1.2]^2 with anti-hourglass coefficient
    0.01.  numpy default_rng(seed), drawn in this order per instance."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ninst):
        n = int(rng.integers(nmin, nmax + 1))
        kind = int(rng.integers(0, 2))
        gamma = float(rng.uniform(1.3, 1.67))
        dr = float(np.exp(rng.uniform(np.log(2.0), np.log(20.0))))
        pr = float(np.exp(rng.uniform(np.log(2.0), np.log(100.0))))
        if kind == abi.INSTANCE_SOD:
            rl, pl, rr, prr = 1.0, 1.0, 1.0 / dr, 1.0 / pr
            out.append(dict(nx=n, ny=n, kind=kind, length=1.0, height=1.0, gamma=gamma,
                            kappa=0.0, rho_l=rl, e_l=pl / ((gamma - 1.0) * rl), rho_r=rr,
                            e_r=prr / ((gamma - 1.0) * rr)))
        else:
            out.append(dict(nx=n, ny=n, kind=kind, length=1.2, height=1.2, gamma=gamma,
                            kappa=0.01, rho_l=1.0, e_l=0.25, rho_r=1.0, e_r=0.0))
    return out


def instances_array(params):
    arr = (abi.Instance * len(params))()
    for a, q in zip(arr, params):
        for k, v in q.items():
            setattr(a, k, v)
    return arr


def ensemble_instance(q, nsteps=200, config=None):
    """The standalone problem of one ensemble instance: the mesh, states and
    tables bl_ensemble_create gives instance q (two regions of q's kappa, one
    material of q's gamma)."""
    nx, ny = q["nx"], q["ny"]
    ielnd, ndx, ndy, indtype = structured_mesh(nx, ny, 0.0, q["length"], 0.0, q["height"])
    nel = nx * ny
    if q["kind"] == abi.INSTANCE_SOD:
        xc, _ = centroids(ielnd, ndx, ndy)
        left = xc < 0.5 * q["length"]
        ielreg = np.where(left, 1, 2).astype(np.int32)
        density = np.where(left, q["rho_l"], q["rho_r"])
        energy = np.where(left, q["e_l"], q["e_r"])
    else:
        ielreg = np.ones(nel, np.int32)
        density = np.full(nel, q["rho_l"])
        energy = np.zeros(nel)
        energy[0] = q["e_l"] / (q["rho_l"] * element_volume(ielnd, ndx, ndy, 0))
    itype, param = ideal_eos([q["gamma"]])
    return Problem(
        name="instance_%s_%dx%d" % ("sod" if q["kind"] == 0 else "sedov", nx, ny),
        ielnd=ielnd, ndx=ndx, ndy=ndy, indtype=indtype, ielreg=ielreg,
        ielmat=np.ones(nel, np.int32), density=density, energy=energy,
        ndu=np.zeros(len(ndx)), ndv=np.zeros(len(ndx)), nmat=1, nreg=2, eos_itype=itype,
        eos_param=param, kappareg=np.array([q["kappa"], q["kappa"]]),
        config=dict(config or ENSEMBLE_CONFIG), nsteps=nsteps, dims=(nx, ny))


def with_relaxation(p, rtype=abi.RELAX_EQUI, nsweep=1, om=1.0, tol=0.0):
    """p with ALE on every step and node relaxation as its mesh motion
    (zexist, not zeul)."""
    p.config["zexist"] = 1
    p.config["zeul"] = 0
    p.config["mintime"] = -np.finfo(np.float64).max
    p.config["maxtime"] = np.finfo(np.float64).max
    p.relax = (int(rtype), int(nsweep), float(om), float(tol))
    p.name = p.name + "_relax%d" % rtype
    return p


def mesh_problem(name, ielnd, ndx, ndy, indtype, gamma=1.4, density=1.0, energy=1.0, **cfg):
    """A uniform ideal-gas state at rest on a given mesh (one region, one
    material)."""
    nel, nnd = len(ielnd) // 4, len(ndx)
    itype, param = ideal_eos([gamma])
    return Problem(name=name, ielnd=np.asarray(ielnd, np.int32), ndx=np.array(ndx, np.float64),
                   ndy=np.array(ndy, np.float64), indtype=np.asarray(indtype, np.int32),
                   ielreg=np.ones(nel, np.int32), ielmat=np.ones(nel, np.int32),
                   density=np.full(nel, density), energy=np.full(nel, energy),
                   ndu=np.zeros(nnd), ndv=np.zeros(nnd), nmat=1, nreg=1, eos_itype=itype,
                   eos_param=param, config=dict(DEFAULT_CONFIG, **cfg))


def harmonic_map(n, eps=0.25, jitter=0.2, seed=7):
    """An n x n mesh of the unit logical square under the inverse of the
    harmonic map xi = x + eps (x^2 - y^2), eta = y.  Returns (ielnd, ndx,
    ndy, indtype, exact_x, exact_y): boundary nodes at their exact positions,
    interior nodes displaced from theirs by U(-jitter, jitter) cell widths,
    and the exact positions of every node.  xi and eta are harmonic in
    (x, y), so the exact positions solve the equipotential (Winslow)
    equations; the grid is not orthogonal (eps != 0)."""
    ielnd, xi, eta, indtype = structured_mesh(n, n, 0.0, 1.0, 0.0, 1.0)
    ex = (-1.0 + np.sqrt(1.0 + 4.0 * eps * (xi + eps * eta * eta))) / (2.0 * eps)
    ey = eta.copy()
    rng = np.random.default_rng(seed)
    inner = indtype > 0
    ndx, ndy = ex.copy(), ey.copy()
    ndx[inner] += rng.uniform(-jitter, jitter, inner.sum()) / n
    ndy[inner] += rng.uniform(-jitter, jitter, inner.sum()) / n
    return ielnd, ndx, ndy, indtype, ex, ey


def sod2mat(nx=2048, ny=256, length=1.0, amplitude=0.02, seed=2024, nsteps=300,
            eulerian=True, jitter=0.0, **cfg):
    """Two-material Sod on [0,L]x[0,L*ny/nx] (square cells): material 1
    (gamma 1.4, rho 1, p 1) left of x = L/2 + A sin(2 pi k y/H + phi), material
    2 (gamma 1.67, rho 0.125, p 0.1) right of it, k and phi drawn from the
    seed.  Each element takes the material of its centroid (pure cells)."""
    height = length * ny / nx
    ielnd, ndx, ndy, indtype = structured_mesh(nx, ny, 0.0, length, 0.0, height,
                                               jitter=jitter, seed=seed)
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, 4))
    phi = float(rng.uniform(0.0, 2.0 * math.pi))
    xc, yc = centroids(ielnd, ndx, ndy)
    left = xc < 0.5 * length + amplitude * np.sin(2.0 * math.pi * k * yc / height + phi)
    nel = nx * ny
    g1, g2 = 1.4, 1.67
    density = np.where(left, 1.0, 0.125)
    energy = np.where(left, 1.0 / ((g1 - 1.0) * 1.0), 0.1 / ((g2 - 1.0) * 0.125))
    itype, param = ideal_eos([g1, g2])
    config = dict(DEFAULT_CONFIG, time_end=1.0e30, **cfg)
    if eulerian:
        config["zeul"] = 1
    mat = np.where(left, 1, 2).astype(np.int32)
    return Problem(
        name="sod2mat_%dx%d" % (nx, ny), ielnd=ielnd, ndx=ndx, ndy=ndy, indtype=indtype,
        ielreg=mat.copy(), ielmat=mat, density=density, energy=energy,
        ndu=np.zeros(len(ndx)), ndv=np.zeros(len(ndx)), nmat=2, nreg=2, eos_itype=itype,
        eos_param=param, config=config, nsteps=nsteps, dims=(nx, ny))


def sod_mm_deck_small():
    """The reference deck input/sod_mm_lag_small: Sod on [0,100]x[0,25],
    100x25, air (gamma 1.4) everywhere with rho=1, e=2.5 from the region IC,
    helium (gamma 1.2, rho=0.125, e=2.0) in the rectangle x > 50.5.  The
    shape test of setup_IC.f90:1585-1634 (intersect: 0.5 when some but not
    all corners are inside; shape_param(5) = 0.5) makes the column
    x in [50,51] mixed with 50/50 volume fractions; its air component takes
    the element's region values (setup_thermodynamic_rationalise,
    setup_IC.f90:740-790), its helium component the material IC."""
    nx, ny = 100, 25
    ielnd, ndx, ndy, indtype = structured_mesh(nx, ny, 0.0, 100.0, 0.0, 25.0)
    nel = nx * ny
    i = np.arange(nel) % nx
    itype, param = ideal_eos([1.4, 1.2])
    mat = np.where(i >= 51, 2, 1).astype(np.int32)
    density = np.where(mat == 2, 0.125, 1.0)
    energy = np.where(mat == 2, 2.0, 2.5)
    config = dict(DEFAULT_CONFIG, time_end=20.0, dt_initial=1.0e-3)
    p = Problem(name="sod_mm_lag_small", ielnd=ielnd, ndx=ndx, ndy=ndy, indtype=indtype,
                ielreg=np.ones(nel, np.int32), ielmat=mat, density=density, energy=energy,
                ndu=np.zeros(len(ndx)), ndv=np.zeros(len(ndx)), nmat=2, nreg=1,
                eos_itype=itype, eos_param=param, config=config, nsteps=100000,
                stop_at_time_end=1, dims=(nx, ny))
    p.set_mixed({int(e): [(1, 0.5, 1.0, 2.5), (2, 0.5, 0.125, 2.0)]
                 for e in np.nonzero(i == 50)[0]})
    return p


def sod2mat_mixed(nx=2048, ny=256, length=1.0, amplitude=0.02, seed=2024, nsteps=300,
                  eulerian=True, sub=8, jitter=0.0, **cfg):
    """sod2mat with mixed interface cells: the volume fraction of material 2
    in each element is the fraction of its sub x sub sample points (bilinear
    in the element) right of the interface; elements with both materials
    present are mixed with components (1, 1-vf) and (2, vf) at the material
    states, and element density and energy are the volume- and
    mass-weighted means."""
    p = sod2mat(nx=nx, ny=ny, length=length, amplitude=amplitude, seed=seed, nsteps=nsteps,
                eulerian=eulerian, jitter=jitter, **cfg)
    height = length * ny / nx
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, 4))
    phi = float(rng.uniform(0.0, 2.0 * math.pi))
    n = p.ielnd.reshape(-1, 4) - 1
    X, Y = p.ndx[n], p.ndy[n]
    t = (np.arange(sub) + 0.5) / sub
    s_, t_ = np.meshgrid(t, t)
    s_, t_ = s_.ravel(), t_.ravel()
    w = np.stack([(1 - s_) * (1 - t_), s_ * (1 - t_), s_ * t_, (1 - s_) * t_], axis=1)
    px, py = X @ w.T, Y @ w.T
    right = px >= 0.5 * length + amplitude * np.sin(2.0 * math.pi * k * py / height + phi)
    vf = right.mean(axis=1)
    g1, g2 = 1.4, 1.67
    st = {1: (1.0, 1.0 / ((g1 - 1.0) * 1.0)), 2: (0.125, 0.1 / ((g2 - 1.0) * 0.125))}
    mixed = np.nonzero((vf > 0.0) & (vf < 1.0))[0]
    cells = {}
    for e in mixed:
        f2 = float(vf[e])
        (r1, e1), (r2, e2) = st[1], st[2]
        cells[int(e)] = [(1, 1.0 - f2, r1, e1), (2, f2, r2, e2)]
        rho = (1.0 - f2) * r1 + f2 * r2
        p.density[e] = rho
        p.energy[e] = ((1.0 - f2) * r1 * e1 + f2 * r2 * e2) / rho
    pure2 = vf >= 1.0
    p.ielmat[:] = np.where(pure2, 2, 1)
    p.ielreg[:] = p.ielmat
    p.density[pure2], p.energy[pure2] = st[2]
    pure1 = vf <= 0.0
    p.density[pure1], p.energy[pure1] = st[1]
    p.set_mixed(cells)
    p.name = "sod2mat_mixed_%dx%d" % (nx, ny)
    return p


# Tait (eos_kn_forms.f90, type 2): p = max(a ((rho/rho0)^b - 1), pmin),
# params (a, b, rho0, pmin); JWL (type 3): p = w rho e + A (1 - w/t1) exp(-t1)
# + B (1 - w/t2) exp(-t2), t1 = R1 rho0/rho, t2 = R2 rho0/rho, params
# (w, A, B, R1, R2, rho0).  Nondimensional water-like and TNT-like values.
TAIT_PARAM = (0.1, 7.0, 1.0, -0.05, 0.0, 0.0)
JWL_PARAM = (0.3, 3.712, 0.0323, 4.15, 0.95, 1.63)


def tait_jwl(nx=40, ny=4, length=1.0, nsteps=60, eulerian=False, **cfg):
    """A JWL detonation-products column (material 1: rho 1.63, e 1) driving a
    Tait liquid (material 2: rho 1, e 0, at its reference density) on
    [0,L]x[0,L*ny/nx]; the interface at mid-length, reflective walls.  The
    reference's two non-ideal EOS forms, whose pow and exp the ideal-gas
    workloads never reach."""
    height = length * ny / nx
    ielnd, ndx, ndy, indtype = structured_mesh(nx, ny, 0.0, length, 0.0, height)
    xc, _ = centroids(ielnd, ndx, ndy)
    left = xc < 0.5 * length
    nel = nx * ny
    itype = np.array([0, 3, 2], np.int32)
    param = np.zeros(abi.NEOSPARAM * 3)
    param[abi.NEOSPARAM:2 * abi.NEOSPARAM] = JWL_PARAM
    param[2 * abi.NEOSPARAM:] = TAIT_PARAM
    mat = np.where(left, 1, 2).astype(np.int32)
    config = dict(DEFAULT_CONFIG, time_end=1.0e30, dt_initial=5.0e-3)
    config.update(cfg)
    if eulerian:
        config["zeul"] = 1
    return Problem(
        name="tait_jwl_%dx%d%s" % (nx, ny, "_eul" if eulerian else ""), ielnd=ielnd, ndx=ndx,
        ndy=ndy, indtype=indtype, ielreg=mat.copy(), ielmat=mat,
        density=np.where(left, JWL_PARAM[5], TAIT_PARAM[2]), energy=np.where(left, 1.0, 0.0),
        ndu=np.zeros(len(ndx)), ndv=np.zeros(len(ndx)), nmat=2, nreg=2, eos_itype=itype,
        eos_param=param, config=config, nsteps=nsteps, dims=(nx, ny))


def noh(n=30, length=1.0, gamma=5.0 / 3.0, nsteps=100, **cfg):
    """Noh implosion on a Cartesian quarter [0,L]^2: rho=1, e=0, radial inward
    velocity of unit speed (the kinematic 'radial' IC of input/noh_small),
    SLIPY bottom, SLIPX left, FREE top/right.  Strong convergence makes the
    divergence control of hydro_kn_getdtdiv active."""
    ielnd, ndx, ndy, indtype = structured_mesh(n, n, 0.0, length, 0.0, length,
                                               bcs=("SLIPY", "FREE", "FREE", "SLIPX"))
    r = np.sqrt(ndx * ndx + ndy * ndy)
    with np.errstate(invalid="ignore", divide="ignore"):
        ndu = np.where(r > 0.0, -ndx / r, 0.0)
        ndv = np.where(r > 0.0, -ndy / r, 0.0)
    nel = n * n
    itype, param = ideal_eos([gamma])
    config = dict(DEFAULT_CONFIG, time_end=0.6, dt_initial=1.0e-4, dt_max=1.0e-2)
    config.update(cfg)
    return Problem(
        name="noh_%dx%d" % (n, n), ielnd=ielnd, ndx=ndx, ndy=ndy, indtype=indtype,
        ielreg=np.ones(nel, np.int32), ielmat=np.ones(nel, np.int32), density=np.ones(nel),
        energy=np.zeros(nel), ndu=ndu, ndv=ndv, nmat=1, nreg=1, eos_itype=itype,
        eos_param=param, config=config, nsteps=nsteps, dims=(n, n))


def taylor_green(n=1024, u0=1.0, p0=10.0, gamma=5.0 / 3.0, nsteps=500, **cfg):
    """A Taylor-Green vortex on [0,1]^2: node velocity u = u0 sin(pi x)
    cos(pi y), v = -u0 cos(pi x) sin(pi y), rho = 1, element pressure
    p0 + (u0^2/4)(cos 2 pi x + cos 2 pi y) at the element centroid (e =
    p/((gamma-1) rho)); reflecting walls (SLIPY bottom/top, SLIPX left/right:
    the field's own normal velocity is zero there).  The velocity differs
    between every pair of neighbouring nodes, so every element is in motion
    from the first step and goes through the limiter every step."""
    ielnd, ndx, ndy, indtype = structured_mesh(n, n, 0.0, 1.0, 0.0, 1.0)
    xc, yc = centroids(ielnd, ndx, ndy)
    pre = p0 + 0.25 * u0 * u0 * (np.cos(2.0 * math.pi * xc) + np.cos(2.0 * math.pi * yc))
    nel = n * n
    itype, param = ideal_eos([gamma])
    config = dict(DEFAULT_CONFIG, time_end=1.0e30)
    config.update(cfg)
    return Problem(
        name="taylor_green_%dx%d" % (n, n), ielnd=ielnd, ndx=ndx, ndy=ndy, indtype=indtype,
        ielreg=np.ones(nel, np.int32), ielmat=np.ones(nel, np.int32), density=np.ones(nel),
        energy=pre / (gamma - 1.0), ndu=u0 * np.sin(math.pi * ndx) * np.cos(math.pi * ndy),
        ndv=-u0 * np.cos(math.pi * ndx) * np.sin(math.pi * ndy), nmat=1, nreg=1,
        eos_itype=itype, eos_param=param, config=config, nsteps=nsteps, dims=(n, n))


def sod_deck_small(eulerian=False):
    """The reference deck input/sod_sm_lag_small (or _eul_small): 100x25 on
    [0,100]x[0,25], dt_initial 1e-3, time_end 20.  The mesh is a uniform
    structured mesh rather than the deck's SOR-generated one."""
    return sod(nx=100, ny=25, length=100.0, height=25.0, time_end=20.0,
               eulerian=eulerian, dt_initial=1.0e-3)
