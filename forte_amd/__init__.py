"""forte_amd: an MI355X (gfx950) FCI solver with Forte's FCISolver interface.

The compute path is the native extension `_forte_amd` (pybind11) over the C ABI library
`_lib/libforte_fci_gfx950.so` (HIP kernels for gfx950). There is no CPU fallback: importing this
package without the built extension raises ImportError, and creating a solver without a gfx950
device raises RuntimeError.
"""
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _forte_amd as _ext
except ImportError as e:  # pragma: no cover - exercised only when the build is missing
    raise ImportError(
        "forte_amd: the native extension is not built; run `python forte_amd/build.py` "
        f"(or __graft_entry__.build()) first ({e})") from e

ActiveSpaceIntegrals = _ext.ActiveSpaceIntegrals
FCISolver = _ext.FCISolver
DavidsonLiuSolver = _ext.DavidsonLiuSolver
RDMs = _ext.RDMs
RDMsType = _ext.RDMsType
seeded_integrals = _ext.seeded_integrals
DeviceArray = _ext.DeviceArray
device_synchronize = _ext.device_synchronize
lp_forward = _ext.lp_forward
lp_loss_grad = _ext.lp_loss_grad
lp_train = _ext.lp_train
LP_NPARAM = _ext.LP_NPARAM

LIBRARY_PATH = os.path.join(_HERE, "_lib", "libforte_fci_gfx950.so")

from .fcidump import chem_to_forte, read_fcidump  # noqa: E402


def make_active_space_integrals(h, eri, scalar_energy=0.0, frozen_core_energy=0.0,
                                nuclear_repulsion_energy=0.0):
    """ActiveSpaceIntegrals from a core Hamiltonian h and chemist-notation ERIs (pq|rs), in
    Forte's restricted layout (forte/integrals/active_space_integrals.cc:578-583)."""
    h = np.ascontiguousarray(h, dtype=np.float64)
    aa, ab = chem_to_forte(np.asarray(eri, dtype=np.float64))
    n = h.shape[0]
    return ActiveSpaceIntegrals(n, h, h, aa, ab, aa, scalar_energy, frozen_core_energy,
                                nuclear_repulsion_energy)


def fci_problem_from_fcidump(fname, frozen_docc=None, restricted_docc=None, active=None, nel=None,
                             multiplicity=None, ms=None, root_sym=None):
    """The active-space FCI problem of an FCIDUMP file under Forte's MO-space options (point-group
    irreps, FROZEN_DOCC / RESTRICTED_DOCC folded into the core, ACTIVE orbitals per irrep):
    dict(ints, na, nb, multiplicity, irrep). See forte_amd.active_space."""
    from .active_space import active_space_from_fcidump
    a = active_space_from_fcidump(fname, frozen_docc, restricted_docc, active, nel, multiplicity, ms, root_sym)
    ints = make_active_space_integrals(a["h"], a["eri"], scalar_energy=a["e_core"])
    ints.set_active_mo_symmetry(a["orbsym"])
    return dict(ints=ints, na=a["na"], nb=a["nb"], multiplicity=a["multiplicity"], irrep=a["irrep"])


def integrals_from_fcidump(fname):
    """(ActiveSpaceIntegrals, na, nb, multiplicity) for an FCIDUMP file (C1, all orbitals active)."""
    d = read_fcidump(fname)
    nel, ms2 = d["nelec"], d["ms2"]
    na = (nel + ms2) // 2
    nb = nel - na
    ints = make_active_space_integrals(d["hcore"], d["eri"], nuclear_repulsion_energy=d["enuc"])
    return ints, na, nb, ms2 + 1


def make_active_space_method(method, na, nb, multiplicity, nroot, as_ints, options=None):
    """Factory in the manner of forte/base_classes/active_space_method.cc:126-184 (FCI only)."""
    if method.upper() != "FCI":
        raise RuntimeError(f"make_active_space_method: only FCI is provided, not {method}")
    s = FCISolver(na, nb, multiplicity, nroot, as_ints)
    if options:
        s.set_options(dict(options))
        if str(options.get("FCI_PRECONDITIONER", "diagonal")).lower() == "learned":
            from .learned import load_weights
            s.set_lp_weights(load_weights(options.get("FCI_LP_WEIGHTS", None) or _default_lp_weights()))
    return s


def _default_lp_weights():
    return os.path.join(_HERE, "data", "lp_weights.npy")


__all__ = [
    "ActiveSpaceIntegrals", "FCISolver", "DavidsonLiuSolver", "RDMs", "RDMsType",
    "seeded_integrals", "DeviceArray", "device_synchronize", "make_active_space_integrals", "integrals_from_fcidump",
    "make_active_space_method", "read_fcidump", "chem_to_forte", "LIBRARY_PATH", "lp_forward", "lp_loss_grad",
    "fci_problem_from_fcidump",
    "lp_train", "LP_NPARAM",
]
