// forte_amd_py.cpp -- pybind11 module _forte_amd: the host side above the C ABI.
//
// Mirrors the Python surface of Forte's FCI path:
//   ActiveSpaceIntegrals-like container   forte/integrals/active_space_integrals.h:41-213
//   FCISolver (ActiveSpaceMethod)         forte/base_classes/active_space_method.h:95-388,
//                                         forte/fci/fci_solver.h:51-246
//   RDMs / RDMsType                       forte/base_classes/rdms.h:116-400, api/rdms_api.cc:49-103
//   DavidsonLiuSolver + test sigma        forte/api/dl_solver_api.cc:45-86
// Status codes from the C ABI become exceptions (RuntimeError in Python), as the reference's
// C++ exceptions do through pybind11.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <cstring>
#include <pybind11/stl.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/forte_fci.h"

namespace py = pybind11;
using arr = py::array_t<double, py::array::c_style | py::array::forcecast>;

namespace {

// The message is read only after the call returned (argument evaluation order is unspecified).
#define check(call, errexpr)                                                          \
    do {                                                                              \
        int st_ = (call);                                                             \
        if (st_ != FCI_OK) throw std::runtime_error(std::string(errexpr));            \
    } while (0)

struct ActiveSpaceIntegrals {
    int n;
    std::vector<double> oei_a, oei_b, tei_aa, tei_ab, tei_bb;
    double scalar_energy = 0.0, frozen_core_energy = 0.0, nuclear_repulsion_energy = 0.0;
    std::vector<int> active_mo_symmetry;  // irreps of the active orbitals (empty: all 0)
    double e_core() const { return scalar_energy + frozen_core_energy + nuclear_repulsion_energy; }
};

std::vector<double> vec_of(const arr& a, size_t expect, const char* what) {
    if (size_t(a.size()) != expect)
        throw std::runtime_error(std::string(what) + ": expected " + std::to_string(expect) + " elements");
    return std::vector<double>(a.data(), a.data() + a.size());
}

py::array_t<double> make_arr(const std::vector<ssize_t>& shape, const std::vector<double>& data) {
    py::array_t<double> a(shape);
    std::copy(data.begin(), data.end(), a.mutable_data());
    return a;
}

enum class RDMsType { spin_dependent, spin_free };

struct RDMs {
    int level = 0, n = 0;
    RDMsType type = RDMsType::spin_dependent;
    std::vector<double> g1a, g1b, g2aa, g2ab, g2bb, g3aaa, g3aab, g3abb, g3bbb;
    py::array_t<double> t6(const std::vector<double>& v) const {
        if (v.empty()) throw std::runtime_error("RDM not available at this level");
        return make_arr({n, n, n, n, n, n}, v);
    }
    // RDMsSpinFree G3 (fci_vector_rdm.cc:127-131): g3aaa + g3aab(pqrstu) + g3aab(prqsut) + g3aab(qrptus)
    //   + g3abb(pqrstu) + g3abb(qprtsu) + g3abb(rpqust) + g3bbb
    std::vector<double> sf_g3() const {
        if (g3aaa.empty()) throw std::runtime_error("3-RDM not available at this level");
        std::vector<double> g(g3aaa.size());
        auto ix = [&](int a, int b, int c, int d, int e, int f) {
            return ((((size_t(a) * n + b) * n + c) * n + d) * n + e) * n + f;
        };
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q)
                for (int r = 0; r < n; ++r)
                    for (int s = 0; s < n; ++s)
                        for (int t = 0; t < n; ++t)
                            for (int u = 0; u < n; ++u) {
                                const size_t i = ix(p, q, r, s, t, u);
                                g[i] = g3aaa[i] + g3aab[i] + g3aab[ix(p, r, q, s, u, t)] + g3aab[ix(q, r, p, t, u, s)] +
                                       g3abb[i] + g3abb[ix(q, p, r, t, s, u)] + g3abb[ix(r, p, q, u, s, t)] + g3bbb[i];
                            }
        return g;
    }
    py::array_t<double> t2(const std::vector<double>& v) const {
        if (v.empty()) throw std::runtime_error("RDM not available at this level");
        return make_arr({n, n}, v);
    }
    py::array_t<double> t4(const std::vector<double>& v) const {
        if (v.empty()) throw std::runtime_error("RDM not available at this level");
        return make_arr({n, n, n, n}, v);
    }
    // RDMsSpinFree: G1 = g1a + g1b; G2 = g2aa + g2ab + g2ab(qpsr) + g2bb (fci_vector_rdm.cc:121-126)
    std::vector<double> sf_g1() const {
        std::vector<double> g(g1a.size());
        for (size_t i = 0; i < g.size(); ++i) g[i] = g1a[i] + g1b[i];
        return g;
    }
    std::vector<double> sf_g2() const {
        if (g2aa.empty()) throw std::runtime_error("2-RDM not available at this level");
        std::vector<double> g(g2aa.size());
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q)
                for (int r = 0; r < n; ++r)
                    for (int s = 0; s < n; ++s) {
                        size_t i = ((size_t(p) * n + q) * n + r) * n + s;
                        size_t j = ((size_t(q) * n + p) * n + s) * n + r;
                        g[i] = g2aa[i] + g2ab[i] + g2ab[j] + g2bb[i];
                    }
        return g;
    }
};

// Device buffer owned by Python (FP64 or FP32), filled and read through host copies.
class DeviceArray {
  public:
    DeviceArray(int64_t n, const std::string& dtype) : n_(n) {
        if (dtype == "float64") es_ = 8;
        else if (dtype == "float32") es_ = 4;
        else throw std::runtime_error("DeviceArray dtype must be float64 or float32");
        dtype_ = dtype;
        check(fci_device_alloc(size_t(n) * es_, &p_), fci_last_error(nullptr));
    }
    ~DeviceArray() {
        if (p_) fci_device_free(p_);
    }
    DeviceArray(const DeviceArray&) = delete;
    void upload(py::array a) {
        if (a.size() != n_) throw std::runtime_error("DeviceArray.upload: size mismatch");
        py::array b = dtype_ == "float64" ? py::array(py::array_t<double, py::array::c_style | py::array::forcecast>(a))
                                          : py::array(py::array_t<float, py::array::c_style | py::array::forcecast>(a));
        check(fci_device_upload(p_, b.data(), size_t(n_) * es_), fci_last_error(nullptr));
    }
    py::array download() const {
        if (es_ == 8) {
            py::array_t<double> a(n_);
            check(fci_device_download(a.mutable_data(), p_, size_t(n_) * es_), fci_last_error(nullptr));
            return a;
        }
        py::array_t<float> a(n_);
        check(fci_device_download(a.mutable_data(), p_, size_t(n_) * es_), fci_last_error(nullptr));
        return a;
    }
    void copy_from(const DeviceArray& o) {
        if (o.n_ != n_ || o.es_ != es_) throw std::runtime_error("DeviceArray.copy_from: shape mismatch");
        check(fci_device_copy(p_, o.p_, size_t(n_) * es_), fci_last_error(nullptr));
    }
    uintptr_t ptr() const { return reinterpret_cast<uintptr_t>(p_); }
    int64_t size() const { return n_; }
    int dtype_code() const { return es_ == 8 ? 0 : 1; }

  private:
    int64_t n_;
    size_t es_ = 8;
    std::string dtype_;
    void* p_ = nullptr;
};

class FCISolver {
  public:
    // FCISolver(StateInfo(na, nb, multiplicity, twice_ms, irrep), nroot, MOSpaceInfo, ActiveSpaceIntegrals)
    // (fci_solver.h:63-64): the active orbitals' irreps come with the integrals
    FCISolver(int na, int nb, int multiplicity, int nroot, std::shared_ptr<ActiveSpaceIntegrals> ints, int irrep = 0)
        : ints_(std::move(ints)) {
        check(fci_create(ints_->n, na, nb, ints_->oei_a.data(), ints_->oei_b.data(), ints_->tei_aa.data(),
                         ints_->tei_ab.data(), ints_->tei_bb.data(), ints_->e_core(), &h_),
              fci_last_error(nullptr));
        fci_default_options(&o_);
        o_.nroot = nroot;
        o_.multiplicity = multiplicity;
        push();
        const auto& os = ints_->active_mo_symmetry;
        bool sym = irrep != 0;
        for (int x : os) sym = sym || x != 0;
        if (sym) {
            std::vector<int> orbsym = os.empty() ? std::vector<int>(ints_->n, 0) : os;
            if (int(orbsym.size()) != ints_->n) throw std::runtime_error("active_mo_symmetry has the wrong length");
            check(fci_set_symmetry(h_, orbsym.data(), irrep), fci_last_error(h_));
        }
    }
    int64_t sector_size() const { return fci_sector_size(h_); }
    py::tuple string_classes(int spin) const {
        const int64_t ns = fci_nstr(h_, spin);
        py::array_t<int> ir(ns);
        py::array_t<int64_t> ad(ns);
        check(fci_string_classes(h_, spin, ir.mutable_data(), ad.mutable_data()), fci_last_error(h_));
        return py::make_tuple(ir, ad);
    }
    py::array_t<double> evec_sector(int root) const {
        py::array_t<double> a(fci_sector_size(h_));
        check(fci_evec_sector(h_, root, a.mutable_data()), fci_last_error(h_));
        return a;
    }
    ~FCISolver() { fci_destroy(h_); }
    FCISolver(const FCISolver&) = delete;

    void push() { check(fci_set_options(h_, &o_), fci_last_error(h_)); }

    // ActiveSpaceMethod / FCISolver setters (active_space_method.h:285-320, fci_solver.h:88-110)
    void set_e_convergence(double v) { o_.e_convergence = v; push(); }
    void set_r_convergence(double v) { o_.r_convergence = v; push(); }
    void set_maxiter(int v) { o_.maxiter = v; push(); }
    void set_die_if_not_converged(bool v) { o_.die_if_not_converged = v; push(); }
    void set_root(int v) { o_.root = v; push(); }
    void set_print(int v) { o_.print = v; push(); }
    void set_ndets_per_guess_state(int v) { o_.ndets_per_guess = v; push(); }
    void set_guess_per_root(int v) { o_.guess_per_root = v; push(); }
    void set_collapse_per_root(int v) { o_.collapse_per_root = v; push(); }
    void set_subspace_per_root(int v) { o_.subspace_per_root = v; push(); }
    void set_precision(const std::string& p) {
        if (p == "fp64") o_.precision = FCI_PREC_FP64;
        else if (p == "mixed") o_.precision = FCI_PREC_MIXED;
        else throw std::runtime_error("precision must be 'fp64' or 'mixed'");
        push();
    }
    void set_mixed_switch(double v) { o_.mixed_switch = v; push(); }
    void set_preconditioner(const std::string& p) {
        if (p == "diagonal") o_.preconditioner = FCI_PRECOND_DIAGONAL;
        else if (p == "learned") o_.preconditioner = FCI_PRECOND_LEARNED;
        else if (p == "olsen") o_.preconditioner = FCI_PRECOND_OLSEN;
        else throw std::runtime_error("preconditioner must be 'diagonal', 'learned' or 'olsen'");
        push();
    }
    // FCISolver::set_options (fci_solver.cc:156-172) from a dict of Forte option names
    void set_options(const py::dict& d) {
        apply_options(d, o_);
        push();
    }
    static void apply_options(const py::dict& d, fci_options& o_) {
        auto set_precision = [&](const std::string& p) {
            if (p == "fp64") o_.precision = FCI_PREC_FP64;
            else if (p == "mixed") o_.precision = FCI_PREC_MIXED;
            else throw std::runtime_error("precision must be 'fp64' or 'mixed'");
        };
        auto set_preconditioner = [&](const std::string& p) {
            if (p == "diagonal") o_.preconditioner = FCI_PRECOND_DIAGONAL;
            else if (p == "learned") o_.preconditioner = FCI_PRECOND_LEARNED;
            else if (p == "olsen") o_.preconditioner = FCI_PRECOND_OLSEN;
            else throw std::runtime_error("preconditioner must be 'diagonal', 'learned' or 'olsen'");
        };
        for (auto item : d) {
            std::string k = py::str(item.first);
            py::handle v = item.second;
            if (k == "DL_MAXITER") o_.maxiter = v.cast<int>();
            else if (k == "E_CONVERGENCE") o_.e_convergence = v.cast<double>();
            else if (k == "R_CONVERGENCE") o_.r_convergence = v.cast<double>();
            else if (k == "ROOT") o_.root = v.cast<int>();
            else if (k == "DL_GUESS_PER_ROOT") o_.guess_per_root = v.cast<int>();
            else if (k == "DL_DETS_PER_GUESS") o_.ndets_per_guess = v.cast<int>();
            else if (k == "DL_COLLAPSE_PER_ROOT") o_.collapse_per_root = v.cast<int>();
            else if (k == "DL_SUBSPACE_PER_ROOT") o_.subspace_per_root = v.cast<int>();
            else if (k == "PRINT") o_.print = v.cast<int>();
            else if (k == "CI_SPIN_ADAPT") o_.spin_adapt = v.cast<bool>() ? 1 : 0;
            else if (k == "CI_SPIN_ADAPT_FULL_PRECONDITIONER") o_.spin_adapt_full_preconditioner = v.cast<bool>() ? 1 : 0;
            else if (k == "FCI_MS0_SYMMETRY") o_.ms0_symmetry = v.cast<bool>() ? 1 : 0;
            else if (k == "FCI_TEST_RDMS") {
                // the brute-force RDM self test is the test suite's job here
            } else if (k == "FCI_PRECISION") set_precision(v.cast<std::string>());
            else if (k == "FCI_PRECONDITIONER") set_preconditioner(v.cast<std::string>());
            else if (k == "FCI_LP_WEIGHTS") {
                // a weights file path, loaded by forte_amd.make_active_space_method
            }
            else throw std::runtime_error("unknown FCI option " + k);
        }
    }
    void set_active_space_integrals(std::shared_ptr<ActiveSpaceIntegrals> ints) {
        ints_ = std::move(ints);
        check(fci_set_integrals(h_, ints_->oei_a.data(), ints_->oei_b.data(), ints_->tei_aa.data(),
                                ints_->tei_ab.data(), ints_->tei_bb.data(), ints_->e_core()),
              fci_last_error(h_));
    }

    double compute_energy() {
        double e = 0.0;
        int st;
        {
            py::gil_scoped_release nogil;
            st = fci_compute_energy(h_, &e);
        }
        check(st, fci_last_error(h_));
        return e;
    }
    std::vector<double> energies() const {
        std::vector<double> e(o_.nroot);
        check(fci_energies(h_, e.data(), nullptr), "compute_energy has not been called");
        return e;
    }
    std::vector<double> spin2() const {
        std::vector<double> s(o_.nroot);
        check(fci_energies(h_, nullptr, s.data()), "compute_energy has not been called");
        return s;
    }
    py::dict stats() const {
        int it, ns, ns64, conv, fb;
        double sec;
        fci_solve_stats(h_, &it, &ns, &ns64, &sec, &conv, &fb);
        py::dict d;
        d["iterations"] = it;
        d["sigma_builds"] = ns;
        d["sigma_builds_fp64"] = ns64;
        d["seconds"] = sec;
        d["converged"] = bool(conv);
        d["fell_back"] = bool(fb);
        int eta = 0;
        fci_solve_ms0(h_, &eta);
        d["ms0_symmetry"] = eta;
        return d;
    }
    py::array_t<double> evec(int root) const {
        int64_t nd = fci_ndet(h_);
        py::array_t<double> a(nd);
        check(fci_evec(h_, root, a.mutable_data()), fci_last_error(h_));
        return a;
    }
    int64_t ndet() const { return fci_ndet(h_); }
    int64_t nstr(int spin) const { return fci_nstr(h_, spin); }

    // GenCI GAS spaces: gas_of_orbital per active orbital, electron bounds per space
    void set_gas(const std::vector<int>& gas_of_orbital, const std::vector<int>& gas_min,
                 const std::vector<int>& gas_max) {
        if (gas_min.size() != gas_max.size()) throw std::runtime_error("set_gas: gas_min and gas_max differ in length");
        if (!gas_min.empty() && gas_of_orbital.size() != size_t(ints_->n))
            throw std::runtime_error("set_gas: one GAS index per active orbital is needed");
        check(fci_set_gas(h_, int(gas_min.size()), gas_of_orbital.data(), gas_min.data(), gas_max.data()),
              fci_last_error(h_));
    }
    int64_t space_size() const { return fci_space_size(h_); }

    // CSF spin adaptation (CI_SPIN_ADAPT; FCISolver::set_spin_adapt, fci_solver.cc:75-79)
    void set_spin_adapt(bool v) { o_.spin_adapt = v ? 1 : 0; push(); }
    void set_spin_adapt_full_preconditioner(bool v) { o_.spin_adapt_full_preconditioner = v ? 1 : 0; push(); }
    // the Ms = 0 spin-flip symmetry of the determinant vectors (not in the reference; on by default)
    void set_ms0_symmetry(bool v) { o_.ms0_symmetry = v ? 1 : 0; push(); }
    std::pair<int64_t, int64_t> csf_size() {
        int64_t nc = 0, nz = 0;
        check(fci_csf_size(h_, &nc, &nz), fci_last_error(h_));
        return {nc, nz};
    }
    int64_t ncsf() { return csf_size().first; }
    py::array_t<double> det_to_csf(py::array_t<double, py::array::c_style | py::array::forcecast> det) {
        if (det.size() != fci_ndet(h_)) throw std::runtime_error("det_to_csf: vector length is not the number of determinants");
        py::array_t<double> out(ncsf());
        check(fci_det_to_csf(h_, det.data(), out.mutable_data()), fci_last_error(h_));
        return out;
    }
    py::array_t<double> csf_to_det(py::array_t<double, py::array::c_style | py::array::forcecast> csf) {
        if (csf.size() != ncsf()) throw std::runtime_error("csf_to_det: vector length is not the number of CSFs");
        py::array_t<double> out(fci_ndet(h_));
        check(fci_csf_to_det(h_, csf.data(), out.mutable_data()), fci_last_error(h_));
        return out;
    }
    py::array_t<double> hdiag_csf() {
        py::array_t<double> out(ncsf());
        check(fci_hdiag_csf(h_, out.mutable_data()), fci_last_error(h_));
        return out;
    }
    py::array_t<double> evec_csf(int root) {
        py::array_t<double> out(ncsf());
        check(fci_evec_csf(h_, root, out.mutable_data()), fci_last_error(h_));
        return out;
    }

    // ActiveSpaceMethod::rdms(root_list, max_rdm_level, type), active_space_method.h:134-136
    std::vector<std::shared_ptr<RDMs>> rdms(const std::vector<std::pair<int, int>>& roots, int level, RDMsType type) {
        std::vector<std::shared_ptr<RDMs>> out;
        const int n = ints_->n;
        const size_t n2 = size_t(n) * n, n4 = n2 * n2;
        for (auto [l, r] : roots) {
            auto R = std::make_shared<RDMs>();
            R->level = level;
            R->n = n;
            R->type = type;
            R->g1a.resize(n2);
            R->g1b.resize(n2);
            if (level >= 2) {
                R->g2aa.resize(n4);
                R->g2ab.resize(n4);
                R->g2bb.resize(n4);
            }
            if (level > 3) throw std::runtime_error("RDMs of level 4 and higher are not implemented");
            int st;
            {
                py::gil_scoped_release nogil;
                st = fci_rdms(h_, l, r, std::min(level, 2), R->g1a.data(), R->g1b.data(),
                              level >= 2 ? R->g2aa.data() : nullptr, level >= 2 ? R->g2ab.data() : nullptr,
                              level >= 2 ? R->g2bb.data() : nullptr);
            }
            check(st, fci_last_error(h_));
            if (level >= 3) {
                const size_t n6 = n4 * n2;
                R->g3aaa.resize(n6);
                R->g3aab.resize(n6);
                R->g3abb.resize(n6);
                R->g3bbb.resize(n6);
                {
                    py::gil_scoped_release nogil;
                    st = fci_rdms3(h_, l, r, R->g3aaa.data(), R->g3aab.data(), R->g3abb.data(), R->g3bbb.data());
                }
                check(st, fci_last_error(h_));
            }
            out.push_back(R);
        }
        return out;
    }

    // ActiveSpaceMethod::transition_rdms(root_list, method2, max_rdm_level, type)
    // (active_space_method.h:151-154; FCISolver: fci_solver_transition_rdms.cc:37-99)
    std::vector<std::shared_ptr<RDMs>> transition_rdms(const std::vector<std::pair<int, int>>& roots,
                                                       const FCISolver& method2, int level, RDMsType type) {
        std::vector<std::shared_ptr<RDMs>> out;
        const int n = ints_->n;
        const size_t n2 = size_t(n) * n;
        for (auto [l, r] : roots) {
            auto R = std::make_shared<RDMs>();
            R->level = level;
            R->n = n;
            R->type = type;
            R->g1a.resize(n2);
            R->g1b.resize(n2);
            int st;
            {
                py::gil_scoped_release nogil;
                st = fci_transition_rdms(h_, l, method2.h_, r, level, R->g1a.data(), R->g1b.data());
            }
            check(st, fci_last_error(h_));
            out.push_back(R);
        }
        return out;
    }

    py::array_t<double> sigma(const arr& C) {
        int64_t nd = fci_ndet(h_);
        if (C.size() != nd) throw std::runtime_error("sigma: vector size must equal the number of determinants");
        py::array_t<double> s(nd);
        int st;
        {
            py::gil_scoped_release nogil;
            st = fci_sigma(h_, C.data(), s.mutable_data());
        }
        check(st, fci_last_error(h_));
        return s;
    }
    void sigma_device(uintptr_t c, uintptr_t s, int dtype, uintptr_t stream) {
        check(fci_sigma_device(h_, reinterpret_cast<const void*>(c), reinterpret_cast<void*>(s), dtype,
                               reinterpret_cast<void*>(stream)),
              fci_last_error(h_));
    }
    py::tuple sigma_bench(const DeviceArray& C, DeviceArray& S, int nrep) {
        if (C.size() != fci_ndet(h_) || S.size() != C.size() || C.dtype_code() != S.dtype_code())
            throw std::runtime_error("sigma_bench: buffers must hold ndet elements of one dtype");
        double ms = 0.0;
        std::vector<double> t(3);
        int st;
        {
            py::gil_scoped_release nogil;
            st = fci_sigma_bench(h_, reinterpret_cast<const void*>(C.ptr()), reinterpret_cast<void*>(S.ptr()),
                                 C.dtype_code(), nrep, &ms, t.data());
        }
        check(st, fci_last_error(h_));
        return py::make_tuple(ms, t);
    }
    void sigma_arrays(const DeviceArray& C, DeviceArray& S) {
        if (C.size() != fci_ndet(h_) || S.size() != C.size() || C.dtype_code() != S.dtype_code())
            throw std::runtime_error("sigma_arrays: buffers must hold ndet elements of one dtype");
        check(fci_sigma_device(h_, reinterpret_cast<const void*>(C.ptr()), reinterpret_cast<void*>(S.ptr()),
                               C.dtype_code(), nullptr),
              fci_last_error(h_));
        check(fci_device_synchronize(), fci_last_error(nullptr));
    }
    void sigma_ms0_arrays(const DeviceArray& C, DeviceArray& S, int eta) {
        if (C.size() != fci_ndet(h_) || S.size() != C.size() || C.dtype_code() != S.dtype_code())
            throw std::runtime_error("sigma_ms0_arrays: buffers must hold ndet elements of one dtype");
        check(fci_sigma_ms0_device(h_, reinterpret_cast<const void*>(C.ptr()), reinterpret_cast<void*>(S.ptr()),
                                   C.dtype_code(), eta, nullptr),
              fci_last_error(h_));
        check(fci_device_synchronize(), fci_last_error(nullptr));
    }
    std::vector<double> sigma_timings() const {
        std::vector<double> t(3);
        fci_sigma_timings(h_, t.data());
        return t;
    }
    double spin2_of(const arr& C) {
        double s = 0.0;
        if (C.size() != fci_ndet(h_)) throw std::runtime_error("spin2_of: wrong vector size");
        check(fci_spin2_of(h_, C.data(), &s), fci_last_error(h_));
        return s;
    }
    // ---- learned preconditioner ----
    void set_lp_weights(py::array_t<float, py::array::c_style | py::array::forcecast> p) {
        if (p.size() != FCI_LP_NPARAM) throw std::runtime_error("expected FCI_LP_NPARAM parameters");
        check(fci_set_lp_weights(h_, p.data()), fci_last_error(h_));
    }
    py::array_t<double> lp_precondition(py::array_t<double, py::array::c_style | py::array::forcecast> r,
                                        double lambda) {
        int64_t nd = fci_ndet(h_);
        if (r.size() != nd) throw std::runtime_error("residual has the wrong length");
        py::array_t<double> out(nd);
        check(fci_lp_precondition(h_, r.data(), lambda, out.mutable_data()), fci_last_error(h_));
        return out;
    }
    py::array_t<float> lp_features(py::array_t<int64_t, py::array::c_style | py::array::forcecast> idx,
                                   double lambda, py::array_t<double, py::array::c_style | py::array::forcecast> r) {
        if (r.size() != fci_ndet(h_)) throw std::runtime_error("residual has the wrong length");
        int64_t m = idx.size();
        py::array_t<float> X({m, int64_t(FCI_LP_NFEAT)});
        check(fci_lp_features(h_, idx.data(), m, lambda, r.data(), X.mutable_data()), fci_last_error(h_));
        return X;
    }
    py::tuple lp_harvest(int64_t per_iter, int64_t max_groups, uint64_t seed) {
        if (per_iter < 1 || max_groups < 1) throw std::runtime_error("per_iter and max_groups must be positive");
        const int64_t ns = per_iter * max_groups;
        std::vector<float> X(size_t(ns) * FCI_LP_NFEAT), d0(ns), e(ns);
        int64_t ng = 0;
        int st;
        {
            py::gil_scoped_release nogil;
            st = fci_lp_harvest(h_, per_iter, max_groups, seed, X.data(), d0.data(), e.data(), &ng);
        }
        check(st, fci_last_error(h_));
        const int64_t m = ng * per_iter;
        py::array_t<float> aX({m, int64_t(FCI_LP_NFEAT)}), ad(m), ae(m);
        std::memcpy(aX.mutable_data(), X.data(), size_t(m) * FCI_LP_NFEAT * sizeof(float));
        std::memcpy(ad.mutable_data(), d0.data(), size_t(m) * sizeof(float));
        std::memcpy(ae.mutable_data(), e.data(), size_t(m) * sizeof(float));
        return py::make_tuple(aX, ad, ae);
    }
    py::array_t<double> hdiag() {
        py::array_t<double> a(fci_ndet(h_));
        check(fci_hdiag(h_, a.mutable_data()), fci_last_error(h_));
        return a;
    }
    py::array_t<uint64_t> strings(int spin) {
        py::array_t<uint64_t> a(fci_nstr(h_, spin));
        check(fci_strings(h_, spin, a.mutable_data()), fci_last_error(h_));
        return a;
    }
    py::tuple list3(int64_t cnt, const std::function<int(double*, uint32_t*, uint32_t*)>& f) {
        if (cnt < 0) throw std::runtime_error("invalid list key");
        py::array_t<double> s(cnt);
        py::array_t<uint32_t> I(cnt), J(cnt);
        check(f(s.mutable_data(), I.mutable_data(), J.mutable_data()), fci_last_error(h_));
        return py::make_tuple(s, I, J);
    }
    py::tuple vo_list(int spin, int p, int q) {
        return list3(fci_vo_size(h_, spin, p, q),
                     [&](double* s, uint32_t* I, uint32_t* J) { return fci_vo_list(h_, spin, p, q, s, I, J); });
    }
    py::tuple oo_list(int spin, int pq) {
        return list3(fci_oo_size(h_, spin, pq),
                     [&](double* s, uint32_t* I, uint32_t* J) { return fci_oo_list(h_, spin, pq, s, I, J); });
    }
    py::tuple vvoo_list(int spin, int p, int q, int r, int t) {
        return list3(fci_vvoo_size(h_, spin, p, q, r, t), [&](double* s, uint32_t* I, uint32_t* J) {
            return fci_vvoo_list(h_, spin, p, q, r, t, s, I, J);
        });
    }
    const fci_options& options() const { return o_; }

  private:
    std::shared_ptr<ActiveSpaceIntegrals> ints_;
    fci_solver* h_ = nullptr;
    fci_options o_;
};

class DavidsonLiuSolver {
  public:
    DavidsonLiuSolver(int64_t size, int nroot, int cpr, int spr) : nroot_(nroot), size_(size) {
        check(fci_dl_create(size, nroot, cpr, spr, &d_), fci_dl_last_error(nullptr));
    }
    ~DavidsonLiuSolver() { fci_dl_destroy(d_); }
    void add_h_diag(const arr& h) { check(fci_dl_add_h_diag(d_, h.data(), h.size()), fci_dl_last_error(d_)); }
    static void flatten(const std::vector<std::vector<std::pair<int64_t, double>>>& v, std::vector<int64_t>& off,
                        std::vector<int64_t>& idx, std::vector<double>& val) {
        off.assign(1, 0);
        for (const auto& x : v) {
            for (auto [i, c] : x) {
                idx.push_back(i);
                val.push_back(c);
            }
            off.push_back(int64_t(idx.size()));
        }
    }
    void add_guesses(const std::vector<std::vector<std::pair<int64_t, double>>>& g) {
        std::vector<int64_t> off, idx;
        std::vector<double> val;
        flatten(g, off, idx, val);
        check(fci_dl_add_guesses(d_, int(g.size()), off.data(), idx.data(), val.data()), fci_dl_last_error(d_));
    }
    void add_project_out_vectors(const std::vector<std::vector<std::pair<int64_t, double>>>& g) {
        std::vector<int64_t> off, idx;
        std::vector<double> val;
        flatten(g, off, idx, val);
        check(fci_dl_add_project_out_vectors(d_, int(g.size()), off.data(), idx.data(), val.data()),
              fci_dl_last_error(d_));
    }
    void add_test_sigma_builder(const arr& M) {
        if (M.size() != size_ * size_) throw std::runtime_error("add_test_sigma_builder: matrix must be size x size");
        check(fci_dl_set_dense_sigma(d_, M.data()), fci_dl_last_error(d_));
    }
    void set_e_convergence(double v) { fci_dl_set_e_convergence(d_, v); }
    void set_r_convergence(double v) { fci_dl_set_r_convergence(d_, v); }
    void set_maxiter(int v) { fci_dl_set_maxiter(d_, v); }
    void set_print_level(int v) { fci_dl_set_print_level(d_, v); }
    void reset() { check(fci_dl_reset(d_), fci_dl_last_error(d_)); }
    bool solve() {
        int c = 0, st;
        {
            py::gil_scoped_release nogil;
            st = fci_dl_solve(d_, &c);
        }
        check(st, fci_dl_last_error(d_));
        return c != 0;
    }
    std::vector<double> eigenvalues() const {
        std::vector<double> e(nroot_);
        fci_dl_eigenvalues(d_, e.data());
        return e;
    }
    py::array_t<double> eigenvector(int k) const {
        py::array_t<double> a(size_);
        check(fci_dl_eigenvector(d_, k, a.mutable_data()), fci_dl_last_error(d_));
        return a;
    }

  private:
    int nroot_;
    int64_t size_;
    fci_dl* d_ = nullptr;
};

}  // namespace

PYBIND11_MODULE(_forte_amd, m) {
    m.doc() = "gfx950 FCI solver (Forte FCISolver drop-in)";
    py::enum_<RDMsType>(m, "RDMsType")
        .value("spin_dependent", RDMsType::spin_dependent)
        .value("spin_free", RDMsType::spin_free);

    py::class_<ActiveSpaceIntegrals, std::shared_ptr<ActiveSpaceIntegrals>>(m, "ActiveSpaceIntegrals")
        .def(py::init([](int n, const arr& oa, const arr& ob, const arr& taa, const arr& tab, const arr& tbb,
                         double scalar, double frozen, double nuclear) {
                 auto a = std::make_shared<ActiveSpaceIntegrals>();
                 size_t n2 = size_t(n) * n, n4 = n2 * n2;
                 a->n = n;
                 a->oei_a = vec_of(oa, n2, "oei_a");
                 a->oei_b = vec_of(ob, n2, "oei_b");
                 a->tei_aa = vec_of(taa, n4, "tei_aa");
                 a->tei_ab = vec_of(tab, n4, "tei_ab");
                 a->tei_bb = vec_of(tbb, n4, "tei_bb");
                 a->scalar_energy = scalar;
                 a->frozen_core_energy = frozen;
                 a->nuclear_repulsion_energy = nuclear;
                 return a;
             }),
             py::arg("nmo"), py::arg("oei_a"), py::arg("oei_b"), py::arg("tei_aa"), py::arg("tei_ab"),
             py::arg("tei_bb"), py::arg("scalar_energy") = 0.0, py::arg("frozen_core_energy") = 0.0,
             py::arg("nuclear_repulsion_energy") = 0.0)
        .def_property_readonly("nmo", [](const ActiveSpaceIntegrals& a) { return a.n; })
        .def("active_mo_symmetry", [](const ActiveSpaceIntegrals& a) {
            return a.active_mo_symmetry.empty() ? std::vector<int>(a.n, 0) : a.active_mo_symmetry;
        })
        .def("set_active_mo_symmetry", [](ActiveSpaceIntegrals& a, const std::vector<int>& s) {
            if (!s.empty() && int(s.size()) != a.n) throw std::runtime_error("one irrep per active orbital");
            a.active_mo_symmetry = s;
        })
        .def("scalar_energy", [](const ActiveSpaceIntegrals& a) { return a.scalar_energy; })
        .def("frozen_core_energy", [](const ActiveSpaceIntegrals& a) { return a.frozen_core_energy; })
        .def("nuclear_repulsion_energy", [](const ActiveSpaceIntegrals& a) { return a.nuclear_repulsion_energy; })
        .def("oei_a", [](const ActiveSpaceIntegrals& a) { return make_arr({a.n, a.n}, a.oei_a); })
        .def("oei_b", [](const ActiveSpaceIntegrals& a) { return make_arr({a.n, a.n}, a.oei_b); })
        .def("tei_aa", [](const ActiveSpaceIntegrals& a) { return make_arr({a.n, a.n, a.n, a.n}, a.tei_aa); })
        .def("tei_ab", [](const ActiveSpaceIntegrals& a) { return make_arr({a.n, a.n, a.n, a.n}, a.tei_ab); })
        .def("tei_bb", [](const ActiveSpaceIntegrals& a) { return make_arr({a.n, a.n, a.n, a.n}, a.tei_bb); });

    py::class_<RDMs, std::shared_ptr<RDMs>>(m, "RDMs")
        .def("max_rdm_level", [](const RDMs& r) { return r.level; })
        .def("rdm_type", [](const RDMs& r) { return r.type; })
        .def("g1a", [](const RDMs& r) { return r.t2(r.g1a); })
        .def("g1b", [](const RDMs& r) { return r.t2(r.g1b); })
        .def("g2aa", [](const RDMs& r) { return r.t4(r.g2aa); })
        .def("g2ab", [](const RDMs& r) { return r.t4(r.g2ab); })
        .def("g2bb", [](const RDMs& r) { return r.t4(r.g2bb); })
        .def("SF_G1", [](const RDMs& r) { return r.t2(r.sf_g1()); })
        .def("SF_G2", [](const RDMs& r) { return r.t4(r.sf_g2()); })
        .def("g3aaa", [](const RDMs& r) { return r.t6(r.g3aaa); })
        .def("g3aab", [](const RDMs& r) { return r.t6(r.g3aab); })
        .def("g3abb", [](const RDMs& r) { return r.t6(r.g3abb); })
        .def("g3bbb", [](const RDMs& r) { return r.t6(r.g3bbb); })
        .def("SF_G3", [](const RDMs& r) { return r.t6(r.sf_g3()); });

    py::class_<FCISolver, std::shared_ptr<FCISolver>>(m, "FCISolver")
        .def(py::init<int, int, int, int, std::shared_ptr<ActiveSpaceIntegrals>, int>(), py::arg("na"), py::arg("nb"),
             py::arg("multiplicity"), py::arg("nroot"), py::arg("as_ints"), py::arg("irrep") = 0)
        .def("sector_size", &FCISolver::sector_size)
        .def("string_classes", &FCISolver::string_classes)
        .def("evec_sector", &FCISolver::evec_sector)
        .def("compute_energy", &FCISolver::compute_energy)
        .def("energies", &FCISolver::energies)
        .def("spin2", &FCISolver::spin2)
        .def("stats", &FCISolver::stats)
        .def("evec", &FCISolver::evec)
        .def("set_gas", &FCISolver::set_gas, py::arg("gas_of_orbital"), py::arg("gas_min"), py::arg("gas_max"))
        .def("space_size", &FCISolver::space_size)
        .def("set_spin_adapt", &FCISolver::set_spin_adapt)
        .def("set_spin_adapt_full_preconditioner", &FCISolver::set_spin_adapt_full_preconditioner)
        .def("set_ms0_symmetry", &FCISolver::set_ms0_symmetry)
        .def("sigma_ms0_arrays", &FCISolver::sigma_ms0_arrays, py::arg("C"), py::arg("S"), py::arg("eta"))
        .def("csf_size", &FCISolver::csf_size)
        .def("ncsf", &FCISolver::ncsf)
        .def("det_to_csf", &FCISolver::det_to_csf)
        .def("csf_to_det", &FCISolver::csf_to_det)
        .def("hdiag_csf", &FCISolver::hdiag_csf)
        .def("evec_csf", &FCISolver::evec_csf)
        .def("ndet", &FCISolver::ndet)
        .def("nstr", &FCISolver::nstr)
        .def("rdms", &FCISolver::rdms, py::arg("root_list"), py::arg("max_rdm_level"),
             py::arg("rdm_type") = RDMsType::spin_dependent)
        .def("transition_rdms", &FCISolver::transition_rdms, py::arg("root_list"), py::arg("method2"),
             py::arg("max_rdm_level"), py::arg("rdm_type") = RDMsType::spin_dependent)
        .def("sigma", &FCISolver::sigma)
        .def("sigma_device", &FCISolver::sigma_device, py::arg("c_ptr"), py::arg("sigma_ptr"), py::arg("dtype"),
             py::arg("stream") = 0)
        .def("sigma_timings", &FCISolver::sigma_timings)
        .def("sigma_bench", &FCISolver::sigma_bench)
        .def("sigma_arrays", &FCISolver::sigma_arrays)
        .def("spin2_of", &FCISolver::spin2_of)
        .def("hdiag", &FCISolver::hdiag)
        .def("set_lp_weights", &FCISolver::set_lp_weights)
        .def("lp_precondition", &FCISolver::lp_precondition)
        .def("lp_features", &FCISolver::lp_features)
        .def("lp_harvest", &FCISolver::lp_harvest, py::arg("per_iter"), py::arg("max_groups"), py::arg("seed") = 0)
        .def("strings", &FCISolver::strings)
        .def("vo_list", &FCISolver::vo_list)
        .def("oo_list", &FCISolver::oo_list)
        .def("vvoo_list", &FCISolver::vvoo_list)
        .def("set_options", &FCISolver::set_options)
        .def("set_active_space_integrals", &FCISolver::set_active_space_integrals)
        .def("set_e_convergence", &FCISolver::set_e_convergence)
        .def("set_r_convergence", &FCISolver::set_r_convergence)
        .def("set_maxiter", &FCISolver::set_maxiter)
        .def("set_die_if_not_converged", &FCISolver::set_die_if_not_converged)
        .def("set_root", &FCISolver::set_root)
        .def("set_print", &FCISolver::set_print)
        .def("set_ndets_per_guess_state", &FCISolver::set_ndets_per_guess_state)
        .def("set_guess_per_root", &FCISolver::set_guess_per_root)
        .def("set_collapse_per_root", &FCISolver::set_collapse_per_root)
        .def("set_subspace_per_root", &FCISolver::set_subspace_per_root)
        .def("set_precision", &FCISolver::set_precision)
        .def("set_mixed_switch", &FCISolver::set_mixed_switch)
        .def("set_preconditioner", &FCISolver::set_preconditioner);

    py::class_<DeviceArray, std::shared_ptr<DeviceArray>>(m, "DeviceArray")
        .def(py::init<int64_t, const std::string&>(), py::arg("n"), py::arg("dtype") = "float64")
        .def("upload", &DeviceArray::upload)
        .def("download", &DeviceArray::download)
        .def("copy_from", &DeviceArray::copy_from)
        .def("ptr", &DeviceArray::ptr)
        .def("size", &DeviceArray::size);

    py::class_<DavidsonLiuSolver, std::shared_ptr<DavidsonLiuSolver>>(m, "DavidsonLiuSolver")
        .def(py::init<int64_t, int, int, int>(), py::arg("size"), py::arg("nroots"), py::arg("collapse_per_root") = 1,
             py::arg("subspace_per_root") = 5)
        .def("add_h_diag", &DavidsonLiuSolver::add_h_diag)
        .def("add_guesses", &DavidsonLiuSolver::add_guesses)
        .def("add_project_out_vectors", &DavidsonLiuSolver::add_project_out_vectors)
        .def("add_test_sigma_builder", &DavidsonLiuSolver::add_test_sigma_builder)
        .def("set_e_convergence", &DavidsonLiuSolver::set_e_convergence)
        .def("set_r_convergence", &DavidsonLiuSolver::set_r_convergence)
        .def("set_maxiter", &DavidsonLiuSolver::set_maxiter)
        .def("set_print_level", &DavidsonLiuSolver::set_print_level)
        .def("reset", &DavidsonLiuSolver::reset)
        .def("solve", &DavidsonLiuSolver::solve)
        .def("eigenvalues", &DavidsonLiuSolver::eigenvalues)
        .def("eigenvector", &DavidsonLiuSolver::eigenvector);

    m.def("device_synchronize", [] {
        py::gil_scoped_release nogil;
        return fci_device_synchronize() == FCI_OK;
    });
    using farr = py::array_t<float, py::array::c_style | py::array::forcecast>;
    m.attr("LP_NPARAM") = FCI_LP_NPARAM;
    m.def("lp_forward", [](farr p, farr X) {
        if (p.size() != FCI_LP_NPARAM || X.ndim() != 2 || X.shape(1) != FCI_LP_NFEAT)
            throw std::runtime_error("lp_forward: bad shapes");
        int64_t m = X.shape(0);
        py::array_t<float> y(m);
        check(fci_lp_forward(p.data(), X.data(), m, y.mutable_data()), fci_last_error(nullptr));
        return y;
    });
    m.def("lp_loss_grad", [](farr p, farr X, farr d0, farr e, float reg) {
        if (p.size() != FCI_LP_NPARAM || X.ndim() != 2 || X.shape(1) != FCI_LP_NFEAT || d0.size() != X.shape(0) ||
            e.size() != X.shape(0))
            throw std::runtime_error("lp_loss_grad: bad shapes");
        float loss = 0.f;
        py::array_t<float> g(FCI_LP_NPARAM);
        check(fci_lp_loss_grad(p.data(), X.data(), d0.data(), e.data(), X.shape(0), reg, &loss, g.mutable_data()),
              fci_last_error(nullptr));
        return py::make_tuple(loss, g);
    }, py::arg("params"), py::arg("X"), py::arg("d0"), py::arg("e"), py::arg("reg") = 0.f);
    m.def("lp_train", [](farr p, farr X, farr d0, farr e, int64_t group, int epochs, float lr, float reg,
                         uint64_t seed) {
        if (p.size() != FCI_LP_NPARAM || X.ndim() != 2 || X.shape(1) != FCI_LP_NFEAT || d0.size() != X.shape(0) ||
            e.size() != X.shape(0) || group < 1 || X.shape(0) % group != 0)
            throw std::runtime_error("lp_train: bad shapes (samples must be a multiple of the group size)");
        py::array_t<float> out(FCI_LP_NPARAM);
        std::memcpy(out.mutable_data(), p.data(), sizeof(float) * FCI_LP_NPARAM);
        py::array_t<float> losses(std::max(epochs, 0));
        int st;
        {
            py::gil_scoped_release nogil;
            st = fci_lp_train(out.mutable_data(), X.data(), d0.data(), e.data(), X.shape(0) / group, group, epochs, lr,
                              reg, seed, losses.mutable_data());
        }
        check(st, fci_last_error(nullptr));
        return py::make_tuple(out, losses);
    }, py::arg("params"), py::arg("X"), py::arg("d0"), py::arg("e"), py::arg("group"), py::arg("epochs"),
       py::arg("lr") = 1e-3f, py::arg("reg") = 1e-4f, py::arg("seed") = 0);
    // many independent problems in one call (fci_batch_solve): problem i = (ints[i], na[i], nb[i])
    m.def("batch_solve", [](const std::vector<std::shared_ptr<ActiveSpaceIntegrals>>& ints, const std::vector<int>& na,
                            const std::vector<int>& nb, int nroot, int multiplicity, const py::dict& options,
                            bool want_evecs) {
        const size_t P = ints.size();
        if (na.size() != P || nb.size() != P) throw std::runtime_error("batch_solve: ints, na and nb differ in length");
        fci_options o;
        fci_default_options(&o);
        o.nroot = nroot;
        o.multiplicity = multiplicity;
        o.die_if_not_converged = 0;
        FCISolver::apply_options(options, o);
        std::vector<int> norb(P), iters(P), conv(P);
        std::vector<const double*> oa(P), ob(P), taa(P), tab(P), tbb(P);
        std::vector<double> ec(P);
        std::vector<py::array_t<double>> ev;
        std::vector<double*> evp(P, nullptr);
        for (size_t i = 0; i < P; ++i) {
            const ActiveSpaceIntegrals& a = *ints[i];
            norb[i] = a.n;
            oa[i] = a.oei_a.data(); ob[i] = a.oei_b.data();
            taa[i] = a.tei_aa.data(); tab[i] = a.tei_ab.data(); tbb[i] = a.tei_bb.data();
            ec[i] = a.e_core();
            if (want_evecs) {
                int64_t ns_a = 1, ns_b = 1;
                for (int t = 0; t < na[i]; ++t) ns_a = ns_a * (a.n - t) / (t + 1);
                for (int t = 0; t < nb[i]; ++t) ns_b = ns_b * (a.n - t) / (t + 1);
                ev.emplace_back(std::vector<py::ssize_t>{nroot, py::ssize_t(ns_a * ns_b)});
                evp[i] = ev.back().mutable_data();
            }
        }
        py::array_t<double> e(std::vector<py::ssize_t>{py::ssize_t(P), nroot});
        fci_batch_stats st;
        std::memset(&st, 0, sizeof(st));
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = fci_batch_solve(int(P), norb.data(), na.data(), nb.data(), oa.data(), ob.data(), taa.data(),
                                 tab.data(), tbb.data(), ec.data(), &o, e.mutable_data(), iters.data(), conv.data(),
                                 want_evecs ? evp.data() : nullptr, &st);
        }
        if (rc != FCI_OK) throw std::runtime_error(fci_batch_last_error());
        py::dict out, stats;
        out["energies"] = e;
        out["iterations"] = iters;
        std::vector<bool> cb(conv.begin(), conv.end());
        out["converged"] = cb;
        if (want_evecs) out["evecs"] = ev;
        stats["setup_seconds"] = st.setup_seconds;
        stats["iterate_seconds"] = st.iterate_seconds;
        stats["converged"] = st.converged;
        stats["iterations"] = st.iterations;
        stats["sigma_builds"] = st.sigma_builds;
        py::list cls;
        for (int k = 0; k < std::min(st.nclasses, FCI_BATCH_MAX_CLASSES); ++k) {
            py::dict c;
            c["norb"] = st.class_norb[k];
            c["problems"] = st.class_nprob[k];
            c["chunk"] = st.class_chunk[k];
            c["subbatch"] = st.class_subbatch[k];
            c["graph_launches"] = st.class_graph_launches[k];
            c["ms_per_iteration"] = st.class_ms_per_iteration[k];
            cls.append(c);
        }
        stats["classes"] = cls;
        out["stats"] = stats;
        return out;
    }, py::arg("ints"), py::arg("na"), py::arg("nb"), py::arg("nroot") = 1, py::arg("multiplicity") = 1,
       py::arg("options") = py::dict(), py::arg("want_evecs") = false);
    m.def("seeded_integrals", [](int n, uint64_t seed) {
        std::vector<double> h(size_t(n) * n), eri(size_t(n) * n * n * n);
        fci_seeded_integrals(n, seed, h.data(), eri.data());
        return py::make_tuple(make_arr({n, n}, h), make_arr({n, n, n, n}, eri));
    });
}
