"""ctypes mirrors of the C-ABI structs and field registry of include/bookleaf_hip.h.

Field ids follow src/dataAPI/dataAPI_id.f90:26-74; the phase aliases follow
dataAPI_id.f90:78-111 (dtstep / lagstep / alestep scratch ids).
"""
import ctypes as C

import numpy as np

SUCCESS, FAILURE = 0, -1
HALT_ALL, HALT_SINGLE = 1, 2
SLEN = 130
NEOSPARAM = 6
# largest mesh of bl_hydro_getdt_noalloc (BL_NOALLOC_MAX_EL)
NOALLOC_MAX_EL = 1 << 24

KIND_R64, KIND_I32, KIND_BOOL = 1, 2, 3
DT_MAXIMUM, DT_GROWTH, DT_INITIAL, DT_CFL, DT_DIV = 1, 2, 3, 4, 5
DT_NAMES = {1: " MAXIMUM", 2: "  GROWTH", 3: " INITIAL", 4: "     CFL", 5: "     DIV"}

FIELDS = """eldensity elenergy elpressure elcs2 elvolume a1 a2 a3 b1 b2 b3 cnwt cnx cny
rscratch21 rscratch22 rscratch23 rscratch24 rscratch25 rscratch26 rscratch27 rscratch28
cnviscx cnviscy elvisc elmass cnmass rscratch11 rscratch12 rscratch13 rscratch14
rscratch15 rscratch16 rscratch17 ndx ndy ndu ndv indtype ielsort1 ielnd ielel ielfc
ielmat ielreg ielsort2 spmass iscratch11 zscratch11 iellocglob indlocglob cpdensity
cpenergy cppressure cpcs2 cpvolume frvolume cpmass frmass cpviscx cpviscy cpvisc cpa1
cpa3 cpb1 cpb3 icpmat icpnext icpprev imxel imxfcp imxncp icpscratch11 icpscratch12
rcpscratch11 rcpscratch21 rcpscratch22 rcpscratch23 rcpscratch24""".split()
ID = {name: i + 1 for i, name in enumerate(FIELDS)}
NFIELD = len(FIELDS)
NAME = {v: k for k, v in ID.items()}

# phase aliases (dataAPI_id.f90:78-111)
DTSTEP = dict(scratch="rscratch11", cnv="rscratch22", ellength="rscratch12",
              cnu="rscratch21", du="rscratch23", dv="rscratch24", dx="rscratch25",
              dy="rscratch26", store="rscratch27")
LAGSTEP = dict(cnu="rscratch21", cnv="rscratch22", cnfx="rscratch23", cnfy="rscratch24",
               elenergy0="rscratch11", ndarea="rscratch12", ndmass="rscratch13",
               ndubar="rscratch14", ndvbar="rscratch15")
ALESTEP = dict(store1="rscratch11", store2="rscratch12", store3="rscratch13",
               store4="rscratch14", store5="rscratch15", store6="rscratch16",
               fcdv="rscratch21", fcdm="rscratch22", flux="rscratch23", rwork1="rscratch24",
               rwork2="rscratch25", rwork3="rscratch26", cnu="rscratch27", cnv="rscratch28",
               indstatus="iscratch11", zactive="zscratch11")


class Sizes(C.Structure):
    _fields_ = [(n, C.c_int32) for n in
                "nel nel1 nel2 nnd nnd1 nnd2 nsz nmat nreg ncp nmx mcp mmx".split()]


class Config(C.Structure):
    _fields_ = [(n, C.c_double) for n in "zcut zerocut dencut accut".split()] + \
               [(n, C.c_double) for n in "cvisc1 cvisc2 cfl_sf div_sf".split()] + \
               [(n, C.c_int32) for n in "zhg zsp ztq pad0".split()] + \
               [(n, C.c_double) for n in "ccut pcut".split()] + \
               [(n, C.c_double) for n in
                "time_start time_end dt_g dt_min dt_max dt_initial".split()] + \
               [(n, C.c_double) for n in "mintime maxtime sf".split()] + \
               [(n, C.c_int32) for n in "adv_type zexist zon zeul".split()]


class Timestep(C.Structure):
    _fields_ = [("time", C.c_double), ("dt", C.c_double), ("dts", C.c_double),
                ("nstep", C.c_int32), ("idtel", C.c_int32), ("zcorrector", C.c_int32),
                ("sdt", C.c_int32)]


class Dt(C.Structure):
    _fields_ = [("rdt", C.c_double), ("idt", C.c_int32), ("sdt", C.c_int32)]


class Error(C.Structure):
    _fields_ = [("ierr", C.c_int32), ("iout", C.c_int32), ("serr", C.c_char * SLEN)]

    @property
    def message(self):
        return bytes(self.serr).decode(errors="replace").strip()


class Instance(C.Structure):
    """bl_instance_t: one instance of an ensemble."""
    _fields_ = [(n, C.c_int32) for n in "nx ny kind pad".split()] + \
               [(n, C.c_double) for n in
                "length height gamma kappa rho_l e_l rho_r e_r".split()]


INSTANCE_SOD, INSTANCE_SEDOV = 0, 1

# node relaxation types (setup_mesh's region types)
RELAX_OFF, RELAX_LIN1, RELAX_LIN2, RELAX_EQUI = 0, 1, 2, 3

# the node-relaxation network: parameter count, inputs, outputs, and the
# offsets of W1 (64,26), b1, W2 (64,64), b2, W3 (2,64), b3 in the flat vector
MLP_NPARAM, MLP_NIN, MLP_NHID, MLP_NOUT, MLP_NNORM = 6018, 26, 64, 2, 56
MLP_OFF = {"W1": 0, "b1": 1664, "W2": 1728, "b2": 5824, "W3": 5888, "b3": 6016}

FLAG_TOTALS = ("volume", "mass", "ie", "ke", "pressure", "pressure_min", "pressure_max",
               "density_min", "density_max")


class Totals(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                "volume mass ie ke pressure_min pressure_max density_min density_max".split()]


class FlagTotals(C.Structure):
    """One row of hydro_dr_shortprint's region or material table."""
    _fields_ = [(n, C.c_double) for n in FLAG_TOTALS]


class Tuning(C.Structure):
    """The launch shape bl_autotune chose."""
    _fields_ = [("limiter_block", C.c_int32), ("limiter_grid", C.c_int32),
                ("ens_chunk", C.c_int32), ("trials", C.c_int32),
                ("best_us", C.c_double), ("default_us", C.c_double),
                ("best_full_us", C.c_double), ("default_full_us", C.c_double)]


def field_shape(sizes, fid):
    """(kind, count) of a dataAPI field, as src/utilities/data/data.f90:169-430."""
    nel, nnd = sizes.nel2, sizes.nnd2
    nsz = max(nel, nnd)
    mcp, mmx = sizes.mcp, sizes.mmx
    n = NAME[fid]
    el = {"eldensity", "elenergy", "elpressure", "elcs2", "elvolume", "a1", "a2", "a3",
          "b1", "b2", "b3", "elvisc", "elmass"}
    cn = {"cnwt", "cnx", "cny", "cnviscx", "cnviscy", "cnmass", "spmass"}
    if n in el:
        return KIND_R64, nel
    if n in cn:
        return KIND_R64, 4 * nel
    if n.startswith("rscratch2"):
        return KIND_R64, 4 * nsz
    if n.startswith("rscratch1"):
        return KIND_R64, nsz
    if n in ("ndx", "ndy", "ndu", "ndv"):
        return KIND_R64, nnd
    if n in ("indtype", "indlocglob"):
        return KIND_I32, nnd
    if n in ("ielsort1", "ielmat", "ielreg", "ielsort2", "iellocglob"):
        return KIND_I32, nel
    if n in ("ielnd", "ielel", "ielfc"):
        return KIND_I32, 4 * nel
    if n == "iscratch11":
        return KIND_I32, nsz
    if n == "zscratch11":
        return KIND_BOOL, nsz
    if n in ("cpviscx", "cpviscy") or n.startswith("rcpscratch2"):
        return KIND_R64, 4 * mcp
    if n.startswith("icp"):
        return KIND_I32, mcp
    if n.startswith("imx"):
        return KIND_I32, mmx
    return KIND_R64, mcp


DTYPE = {KIND_R64: np.float64, KIND_I32: np.int32, KIND_BOOL: np.uint8}


def declare(lib, prefix):
    """Declare the argument types of the bl_*-style entry points of a library
    whose functions are named <prefix>_<name> (bl for the HIP library, orc for
    the oracle)."""
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "init_state": [vp], "hydro_init": [vp],
        "hydro_getdt": [vp, P(Dt), P(Dt), P(Error)],
        "hydro_setpredictor": [vp], "hydro_setcorrector": [vp],
        "hydro_getforce": [vp, P(Timestep)],
        "hydro_getstate": [vp, P(Timestep), P(Error)],
        "hydro_getenergy": [vp, P(Timestep)], "hydro_geteos": [vp],
        "hydro_initacceleration": [vp], "hydro_scatteracceleration": [vp],
        "hydro_getacceleration": [vp], "hydro_applyacceleration": [vp, P(Timestep)],
        "geometry_getgeometry": [vp, P(Error)], "density_getdensity": [vp],
        "ale_active": [vp, P(Timestep)], "ale_getmeshstatus": [vp],
        "ale_getfluxvolume": [vp, P(Timestep)], "ale_advect": [vp, P(Timestep), P(Error)],
        "ale_hydroupdate": [vp, P(Error)],
        "run": [vp, C.c_int32, C.c_int32, P(Timestep), P(C.c_double), P(C.c_int32),
                P(C.c_int32), P(C.c_int32), P(Error)],
        "totals": [vp, P(Totals)],
        "shortprint": [vp, P(FlagTotals), P(FlagTotals)],
        "sod_l1": [vp, C.c_double, C.c_double, C.c_double, P(C.c_double)],
        "set_diagnostics": [vp, C.c_int32, C.c_double, C.c_double, C.c_int32],
        "diagnostics": [vp, C.c_int32, P(C.c_double), P(C.c_int32)],
        "set_relaxation": [vp, C.c_int32, C.c_int32, C.c_double, C.c_double],
        "relax_mesh": [vp, C.c_int32, C.c_int32, C.c_double, C.c_double],
    }
    for name, args in sig.items():
        f = getattr(lib, prefix + "_" + name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = C.c_int
    return lib
