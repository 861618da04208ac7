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


def gas_layout(gas_per_irrep):
    """Active orbitals of GAS spaces given per irrep, as Forte's GAS1..GAS6 options hold them
    (e.g. [[2, 0, 1, 1], [1, 0, 0, 1]]), in MOSpaceInfo's order: irrep-major, and within an irrep
    GAS1 orbitals first. Returns (orbsym, gas_of_orbital), one entry per active orbital."""
    nirrep = max(len(g) for g in gas_per_irrep)
    orbsym, gas_of = [], []
    for h in range(nirrep):
        for g, counts in enumerate(gas_per_irrep):
            m = counts[h] if h < len(counts) else 0
            orbsym += [h] * m
            gas_of += [g] * m
    return orbsym, gas_of


def _gas_options(options, nel):
    """GASn / GASnMIN / GASnMAX (forte/options.yaml:1632-1680; first entry of a per-state list)."""
    spaces, gmin, gmax = [], [], []
    for k in range(1, 7):
        if f"GAS{k}" not in options:
            continue
        counts = list(options[f"GAS{k}"])
        size = sum(counts)
        if size == 0:
            continue
        lo, hi = options.get(f"GAS{k}MIN", [0]), options.get(f"GAS{k}MAX", [min(2 * size, nel)])
        lo = lo[0] if isinstance(lo, (list, tuple)) else lo
        hi = hi[0] if isinstance(hi, (list, tuple)) else hi
        spaces.append(counts)
        gmin.append(int(lo))
        gmax.append(int(min(hi, 2 * size, nel)))
    return spaces, gmin, gmax


def make_active_space_method(method, na, nb, multiplicity, nroot, as_ints, options=None):
    """Factory in the manner of forte/base_classes/active_space_method.cc:126-184: "FCI", and
    "GENCI" with the GASn / GASnMIN / GASnMAX options (the integrals' orbitals in the order of
    gas_layout)."""
    m = method.upper()
    if m not in ("FCI", "GENCI"):
        raise RuntimeError(f"make_active_space_method: FCI and GENCI are provided, not {method}")
    options = dict(options or {})
    gas_keys = [k for k in options if k.startswith("GAS")]
    gas = {k: options.pop(k) for k in gas_keys}
    s = FCISolver(na, nb, multiplicity, nroot, as_ints)
    if m == "GENCI" and gas:
        spaces, gmin, gmax = _gas_options(gas, na + nb)
        _, gas_of = gas_layout(spaces)
        s.set_gas(gas_of, gmin, gmax)
    if options:
        s.set_options(dict(options))
        if str(options.get("FCI_PRECONDITIONER", "diagonal")).lower() == "learned":
            from .learned import load_weights
            s.set_lp_weights(load_weights(options.get("FCI_LP_WEIGHTS", None) or _default_lp_weights()))
    return s


def compute_average_rdms(solver, weights, max_rdm_level=2, rdm_type=None):
    """Weighted average of the roots' RDMs, sum_r w_r RDMs(r, r) (ActiveSpaceSolver::
    compute_average_rdms, forte/base_classes/active_space_solver.cc, for one solver's roots; the
    MCSCF driver's call after each compute_energy, mcscf/mcscf_2step.cc:187-198). Returns the
    spin-dependent blocks {"g1a", "g1b"[, "g2aa", "g2ab", "g2bb"]} or, with RDMsType.spin_free,
    {"G1"[, "G2"]}."""
    rdm_type = RDMsType.spin_dependent if rdm_type is None else rdm_type
    out = {}
    for r, w in enumerate(weights):
        if w == 0.0:
            continue
        rd = solver.rdms([(r, r)], max_rdm_level, rdm_type)[0]
        if rdm_type == RDMsType.spin_free:
            blocks = {"G1": rd.SF_G1()}
            if max_rdm_level >= 2:
                blocks["G2"] = rd.SF_G2()
        else:
            blocks = {"g1a": rd.g1a(), "g1b": rd.g1b()}
            if max_rdm_level >= 2:
                blocks.update(g2aa=rd.g2aa(), g2ab=rd.g2ab(), g2bb=rd.g2bb())
        for k, v in blocks.items():
            out[k] = out[k] + w * v if k in out else w * v
    return out


def _default_lp_weights():
    return os.path.join(_HERE, "data", "lp_weights.npy")


__all__ = [
    "ActiveSpaceIntegrals", "FCISolver", "DavidsonLiuSolver", "RDMs", "RDMsType",
    "seeded_integrals", "DeviceArray", "device_synchronize", "make_active_space_integrals", "integrals_from_fcidump",
    "make_active_space_method", "read_fcidump", "chem_to_forte", "LIBRARY_PATH", "lp_forward", "lp_loss_grad",
    "fci_problem_from_fcidump", "gas_layout", "compute_average_rdms",
    "lp_train", "LP_NPARAM",
]
