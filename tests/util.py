"""Shared helpers of the test suite: system builders and comparisons."""
import os

import numpy as np

from mspms2_amd.system import system_of  # noqa: F401  (re-exported for the tests)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

ENERGY_KEYS = ("uvdw", "ushift", "ureal", "uexcl", "ufourier", "uself", "uelec", "usflj",
               "uinter", "virial_inter")
SHORT_KEYS = ("ubond", "uangle", "udih", "uimp", "uintra", "virial_intra")


def tube_grid_system(hx=0.5, natom=300):
    """The 300-site methane-in-tube system with the tube given as a
    tricubic Hermite grid (SF_TASOS) tabulated from the closed form."""
    from mspms2_amd import configs, sfgrid
    rd = configs.methane_tube(natom=natom, length=60.0, box_xy=40.0, tube_r=15.0, rc=10.0)
    s = system_of(rd)
    s.params["iSF_type"] = 3
    n = int(round(40.0 / hx)) + 1
    s.grid = sfgrid.tube_grid([(3.73, 148.0)], R=15.0, x0=-20.0, y0=-20.0, nx=n, ny=n, hx=hx,
                              hy=hx, nz=5, hz=15.0, sigma_s=3.37, eps_s=27.68)
    return s


def rel_rms(f, g):
    """||f - g|| / ||g|| over all atoms and components."""
    g = np.asarray(g)
    d = np.linalg.norm(np.asarray(f) - g)
    n = np.linalg.norm(g)
    return d / n if n > 0 else d


def long_energy_floor(t_ref):
    """Nonbonded energy scale of a system (frclong terms)."""
    return energy_scale(t_ref) + abs(t_ref["uvdw"])


def force_close(f, g, t_ref, natom, rc, rtol):
    """||f - g|| <= rtol * max(||g||, force scale).  The floor is the
    per-atom force implied by the virial and LJ energy (|W| + |U|)/(N rc):
    on a perfect lattice the reference forces are rounding noise (~1e-14)
    and a relative error means nothing."""
    g = np.asarray(g)
    d = np.linalg.norm(np.asarray(f) - g)
    fs = (abs(t_ref["virial_inter"]) + abs(t_ref["uvdw"])) / (natom * rc) * np.sqrt(g.size)
    return d <= rtol * max(np.linalg.norm(g), fs), d / max(np.linalg.norm(g), fs)


def energy_scale(t):
    """Scale of the electrostatic cancellation uelec = ureal + ufourier -
    uself - uexcl (+ uvacuum)."""
    return sum(abs(t.get(k, 0.0)) for k in ("ureal", "ufourier", "uself", "uexcl", "uvacuum"))


def assert_long_close(t, ref, rtol, what=""):
    """Each energy within rtol of the reference, relative to its own size
    (uelec / uinter / virial relative to the cancelling components)."""
    scale_el = energy_scale(ref)
    for k in ENERGY_KEYS:
        a, b = t[k], ref[k]
        s = abs(b)
        if k in ("uelec", "uinter", "virial_inter"):
            s = max(s, scale_el, abs(ref["uvdw"]))
        assert abs(a - b) <= rtol * max(s, 1e-300) + 1e-300, \
            f"{what} {k}: {a!r} vs {b!r} (rel {abs(a - b) / max(s, 1e-300):.3e}, tol {rtol})"


def assert_short_close(t, ref, rtol, what="", floor=0.0):
    """frcshort terms within rtol of max(|term|, |uintra|, floor).  Pass the
    system's nonbonded energy scale as floor where the geometry sits at the
    bonded minima (rigid-template inputs): there the reference's bonded
    energies are rounding noise (~1e-24) and a relative error means
    nothing."""
    for k in SHORT_KEYS:
        a, b = t[k], ref[k]
        s = max(abs(b), abs(ref["uintra"]), floor, 1e-300)
        assert abs(a - b) <= rtol * s, f"{what} {k}: {a!r} vs {b!r}"
