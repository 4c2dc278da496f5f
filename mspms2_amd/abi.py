"""ctypes mirror of include/mspms_gpu.h (structs and constants).

The layouts here must match the C header field for field; tests/test_abi.py
checks the sizes against the compiled library.
"""
import ctypes as C

MSPMS_OK = 0
MSPMS_EINVAL = -22
MSPMS_EDOM = -33
MSPMS_ERANGE = -34
MSPMS_ENOMEM = -12
MSPMS_EIO = -5
MSPMS_ENOTFINITE = -74

ERROR_NAMES = {
    MSPMS_EINVAL: "EINVAL", MSPMS_EDOM: "EDOM", MSPMS_ERANGE: "ERANGE",
    MSPMS_ENOMEM: "ENOMEM", MSPMS_EIO: "EIO", MSPMS_ENOTFINITE: "ENOTFINITE",
}

GHOST_NONE, GHOST_LJ, GHOST_FULL = 0, 1, 2
BOND_NONE, BOND_HARMONIC, BOND_MORSE, BOND_FENE = 0, 1, 2, 3
ANGLE_NONE, ANGLE_HARMONIC, ANGLE_TR_WATER = 0, 1, 2
DIH_NONE, DIH_OPLS_COSIN, DIH_CHARMM = 0, 1, 2
IMP_NONE, IMP_CHARMM = 0, 1
SF_NONE, SF_HYPERGEO, SF_ATOM_EXPLICIT, SF_TASOS, SF_MY_INTERP = 0, 1, 2, 3, 4
ELEC_NONE, ELEC_EWALD, ELEC_WOLF, ELEC_SIMPLE_COULOMB = 0, 1, 2, 3
EWALD_BC_TINFOIL, EWALD_BC_VACUUM = 1, 2
PREC_FP64, PREC_MIXED = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)
_fp = C.POINTER(C.c_float)


class Species(C.Structure):
    _fields_ = [
        ("natom", C.c_int), ("nmole", C.c_int),
        ("ghost_type", _ip), ("interp_type", _ip),
        ("aw", _dp), ("sigma", _dp), ("epsilon", _dp), ("charge", _dp),
        ("nbond", C.c_int), ("bnd_idx", _ip), ("bnd_type", _ip), ("bnd_para", _dp),
        ("nangle", C.c_int), ("agl_idx", _ip), ("agl_type", _ip), ("agl_para", _dp),
        ("agl_unique", _u8p),
        ("ndih", C.c_int), ("dih_idx", _ip), ("dih_type", _ip), ("dih_para", _dp),
        ("dih_unique", _u8p),
        ("nimp", C.c_int), ("imp_idx", _ip), ("imp_type", _ip), ("imp_para", _dp),
        ("nnbp", C.c_int),
    ]


class SFGrid(C.Structure):
    _fields_ = [
        ("ntype", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
        ("nodes", _fp), ("zperiod", C.c_double),
    ]


class SFCubes(C.Structure):
    """mspms_sf_cubes: myinterp cube polynomials (init_my_interp's data)."""
    _fields_ = [
        ("ntype", C.c_int), ("ncx", C.c_int), ("ncy", C.c_int), ("ncz", C.c_int),
        ("xmin", C.c_double), ("ymin", C.c_double), ("zmax", C.c_double),
        ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
        ("tables", C.POINTER(C.POINTER(C.c_double))),
    ]


class Params(C.Structure):
    _fields_ = [
        ("sigma_base", C.c_double), ("epsilon_base", C.c_double),
        ("boxlx", C.c_double), ("boxly", C.c_double), ("boxlz", C.c_double),
        ("rcuton", C.c_double), ("rcutoff", C.c_double), ("rcutoffelec", C.c_double),
        ("isLJswitchOn", C.c_int), ("f0", C.c_double), ("iChargeType", C.c_int),
        ("kappa", C.c_double),
        ("kmaxx", C.c_int), ("kmaxy", C.c_int), ("kmaxz", C.c_int), ("ksqmax", C.c_int),
        ("fEwald_BC", C.c_int), ("iSF_type", C.c_int), ("iExternal_FF_type", C.c_int),
        ("rcutoffsq", C.c_double), ("rcutonsq", C.c_double), ("rcutoffelecsq", C.c_double),
        ("roff2_minus_ron2_cube", C.c_double), ("shift1", C.c_double), ("shift4", C.c_double),
        ("coulomb_prefactor", C.c_double),
        ("twopi_lx", C.c_double), ("twopi_ly", C.c_double), ("twopi_lz", C.c_double),
        ("bfactor_ewald", C.c_double), ("vfactor_ewald", C.c_double),
        ("twopi_over_3v", C.c_double),
        ("wolfvcon1", C.c_double), ("wolfvcon2", C.c_double), ("wolfvcon3", C.c_double),
        ("wolffcon1", C.c_double), ("wolffcon2", C.c_double),
        ("ntube", C.c_int), ("tube_x", _dp), ("tube_y", _dp), ("tube_r", _dp),
        ("const_3pisq_theta", C.c_double), ("solid_sigma", C.c_double),
        ("solid_epsilon", C.c_double),
        ("solid_natom", C.c_int), ("solid_xx", _dp), ("solid_yy", _dp), ("solid_zz", _dp),
        ("grid", C.POINTER(SFGrid)),
        ("cubes", C.POINTER(SFCubes)),
    ]


class LongTerms(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "uvdw", "ushift", "ureal", "uexcl", "ufourier", "uself", "uvacuum", "uelec",
        "usflj", "uotherff", "uinter", "virial_inter")] + [
        ("n_lj_pairs", C.c_int64), ("n_elec_pairs", C.c_int64),
        ("lj_pair_hash", C.c_uint64), ("elec_pair_hash", C.c_uint64)]

    ENERGY_FIELDS = ("uvdw", "ushift", "ureal", "uexcl", "ufourier", "uself", "uvacuum",
                     "uelec", "usflj", "uotherff", "uinter", "virial_inter")

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ShortTerms(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "ubond", "uangle", "udih", "uimp", "uintra", "virial_intra", "usflj")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ParamGrad(C.Structure):
    _fields_ = [("dE_depsilon", _dp), ("dE_dsigma", _dp), ("dE_dcharge", _dp)]


TUNE_MAX = 16


class TuneResult(C.Structure):
    _fields_ = [("cell_atoms", C.c_double), ("waves_per_cluster", C.c_int),
                ("ncand", C.c_int), ("ms", C.c_double),
                ("cand_cell_atoms", C.c_double * TUNE_MAX), ("cand_wpc", C.c_int * TUNE_MAX),
                ("cand_ms", C.c_double * TUNE_MAX)]

    def as_dict(self):
        n = self.ncand
        return {"cell_atoms": self.cell_atoms, "waves_per_cluster": self.waves_per_cluster,
                "ms": self.ms,
                "candidates": [(self.cand_cell_atoms[k], self.cand_wpc[k], self.cand_ms[k])
                               for k in range(n)]}


class NvtState(C.Structure):
    """mspms_nvt_state: one replica's Nose-Hoover thermostats (vver_nh_3)."""
    _fields_ = [(k, C.c_double) for k in ("qq", "qqs", "gg", "ggs", "treq", "ss", "ps", "sss",
                                          "pss", "unhts", "unhtss", "ukin", "tinst")] + \
        [("nfree", C.c_int), ("pad", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class NptState(C.Structure):
    """mspms_npt_state: one replica's NPT thermostat / barostat (npt_respa)."""
    _fields_ = [(k, C.c_double) for k in ("Qts", "Qbs", "treq", "preq", "vts", "vbs", "rts",
                                          "pljlrc", "utsbs", "ukin", "tinst", "pinst")] + \
        [("boxl", C.c_double * 3), ("boxv", C.c_double), ("nfree", C.c_int),
         ("lj_tail", C.c_int)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "boxl"}
        d["boxl"] = list(self.boxl)
        return d


def dptr(a):
    """double* of a contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return _dp()
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return _ip()
    return a.ctypes.data_as(_ip)


def u8ptr(a):
    if a is None:
        return _u8p()
    return a.ctypes.data_as(_u8p)


def fptr(a):
    if a is None:
        return _fp()
    return a.ctypes.data_as(_fp)


class SfGridGenSpec(C.Structure):
    """mspms_sfgridgen_spec: the grid generator's input (tools/sfgridgen.c:104-117)."""
    _fields_ = [("f_sigma", C.c_double), ("f_epsilon", C.c_double), ("grid_itvl", C.c_double),
                ("rcuton", C.c_double), ("rcutoff", C.c_double), ("isLJswitch", C.c_int),
                ("s_natom", C.c_int), ("uclx", C.c_double), ("ucly", C.c_double),
                ("uclz", C.c_double), ("s_sigma", C.c_double), ("s_epsilon", C.c_double),
                ("sx", C.POINTER(C.c_double)), ("sy", C.POINTER(C.c_double)),
                ("sz", C.POINTER(C.c_double))]
