// C ABI and the calculator object behind it: the pre/post-processing of
// HansenCoppensStructureFactorCalculator::calculateStructureFactorsAndDerivatives
// (src/Scattering/HansenCoppensStructureFactorCalculator.cpp:380-508) around the HIP engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <thread>

#include "../../include/discamb_amd.h"
#include "hc_engine.hpp"
#include "hc_internal.hpp"
#include "hc_lcs.hpp"

namespace discamb_amd {

static thread_local std::string g_last_error;

// Page-locked host memory (hipHostMalloc), grown on demand.
struct PinnedBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n)
    {
        if (n <= bytes && p) return;
        release();
        if (hipHostMalloc(&p, std::max<size_t>(n, 256), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            throw Error(DH_E_MEMORY, "pinned host allocation failed");
        }
        bytes = std::max<size_t>(n, 256);
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~PinnedBuffer() { release(); }
    double* d() const { return static_cast<double*>(p); }
};

// f(a0, a1) over [0, n) split across up to 16 host threads, at least min_each atoms per thread
// (a thread start costs about as much as a few hundred frame derivatives)
template <typename F>
static void parallel_atoms(int n, int min_each, F&& f)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const int nt = std::max(1, std::min({hw > 0 ? (int)hw : 1, n / std::max(min_each, 1), 16}));
    std::vector<std::exception_ptr> err(nt);
    auto run = [&](int t) {
        try {
            f((int)((long)n * t / nt), (int)((long)n * (t + 1) / nt));
        } catch (...) {
            err[t] = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(run, t);
    run(0);
    for (std::thread& th : pool) th.join();
    for (const std::exception_ptr& e : err)
        if (e) std::rethrow_exception(e);
}

struct Calculator {
    Cell cell;
    SpaceGroupInfo sg;
    EngineModel model;
    std::vector<double> mult_w;      // multiplicity / |G|
    std::vector<int> atom_type, atom_wfn;
    std::vector<int> eng_atom_type;  // engine type per atom (atom_type unless split by core wavefunction)
    int xyz_conv = DH_XYZ_FRACTIONAL, adp_conv = DH_ADP_U_CIF, d_xyz_conv = DH_XYZ_FRACTIONAL,
        d_adp_conv = DH_ADP_U_CIF;
    bool centric_origin = false;
    Precision prec = kFp64;
    std::unique_ptr<Engine> engine;
    // independent-atom model ("CPU_IAM"): its own engine (Gaussian form factors, inversion
    // partners explicit with f'' unconjugated, as HansenCoppens_SF_Engine::calculateSF_IAM)
    bool iam = false;
    std::unique_ptr<Engine> iam_engine;
    std::vector<int> iam_atom_type;
    std::string iam_table = "Waasmeier-Kirfel";
    std::vector<std::string> wfn_label;
    std::vector<int> wfn_z;
    std::vector<std::complex<double>> wfn_anom;
    int device = 0;
    bool electron = false;
    // local frames recomputed from the positions at every upload (and the chain rule through them)
    std::vector<LcsDef> lcs_defs;
    bool lcs_chain = false;
    std::vector<double> last_lcs, last_xyz_frac;
    // host staging of upload(), reused between calls; centring multipliers of the last reflection list
    EngineAtoms up_atoms;
    std::vector<double> up_dtdf;
    std::vector<int> mult_hkl;
    std::vector<double> mult_k;
    Engine& eng() { return iam ? *iam_engine : *engine; }
    bool eng_centric_origin() const { return !iam && centric_origin; }
    // Groups whose operations go beyond sign matrices but contain a diagonal subgroup H of two or
    // four operations (tetragonal, hexagonal, cubic): the summed calls (host and device-resident)
    // run on cos_engine, built with the operations of H, over the images g(x) of every atom under
    // the coset representatives cos_reps (coset_plan), and fold the images' derivatives back;
    // batched steps take it too.  Per-reflection rows and form factors keep the main engine and
    // the full operation list.
    std::unique_ptr<Engine> cos_engine;
    std::vector<EngineOp> cos_reps;
    EngineAtoms cos_atoms;
    std::vector<double> cos_rows;
    Engine* last_run = nullptr;       // the engine of the last run (dh_hc_last_timing)
    Engine& host_eng() { return iam ? *iam_engine : (cos_engine && !engine->tune.general_path) ? *cos_engine : *engine; }
    bool cosets() { return !iam && &host_eng() == cos_engine.get(); }
    // Every engine of the calculator drops its inputs: after an upload or a new reflection list
    // only the engine that took it may run (which engine serves a call changes with the kernel
    // options and the IAM table, and another engine's older inputs could differ in size).
    void invalidate_all()
    {
        for (Engine* e : {engine.get(), cos_engine.get(), iam_engine.get()})
            if (e) e->invalidate_inputs();
    }
    // state of the last upload
    int n_atoms = 0, n_hkl = 0;
    std::vector<int> adp_size;
    std::vector<double> occ;
    std::vector<double> d10;
    unsigned last_selector = 0;
    std::vector<AtomType> types;
    // per-type multipole derivatives of the last run with DH_D_MULTIPOLE
    std::vector<double> d_plm, d_pval, d_kappa, d_kappa_prime;
    bool have_mp = false;
    // per-reflection rows: device blocks and pinned host staging of dh_hc_per_hkl / _rows
    bool rows_uploaded = false;
    DeviceBuffer rows_dev[2], f_dev[2];
    PinnedBuffer rows_host[2], f_host[2];
    hipEvent_t rows_ev[2] = {nullptr, nullptr}, ev_in = nullptr, ev_out = nullptr;
    ~Calculator()
    {
        for (hipEvent_t e : {rows_ev[0], rows_ev[1], ev_in, ev_out})
            if (e) (void)hipEventDestroy(e);
    }
};

// explicit operations; with expand_centre the inversion partners (-R, -t + t_inv) are listed
// too, their anomalous term conjugated (partner_anom = -1, Engine3) or not (+1, Engine v1)
// A rotation that only flips coordinate axes, fractional and Cartesian alike: its sign pattern
// (bit i: axis i flips), or -1.
static int sign_pattern(const EngineOp& o)
{
    int pat = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const int f = o.rf[3 * i + j];
            const double c = o.rc[3 * i + j];
            if (i != j) {
                if (f != 0 || std::fabs(c) > 1e-12) return -1;
            } else {
                if ((f != 1 && f != -1) || std::fabs(c - f) > 1e-12) return -1;
                if (f == -1) pat |= 1 << i;
            }
        }
    return pat;
}

// h o g: x -> R_h (R_g x + t_g) + t_h
static EngineOp compose_ops(const EngineOp& h, const EngineOp& g)
{
    EngineOp o;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            int r = 0;
            double c = 0.0;
            for (int k = 0; k < 3; k++) {
                r += h.rf[3 * i + k] * g.rf[3 * k + j];
                c += h.rc[3 * i + k] * g.rc[3 * k + j];
            }
            o.rf[3 * i + j] = r;
            o.rc[3 * i + j] = c;
        }
        double t = h.tf[i];
        for (int k = 0; k < 3; k++) t += h.rf[3 * i + k] * g.tf[k];
        o.tf[i] = t;
    }
    o.anom_sign = h.anom_sign * g.anom_sign;
    return o;
}

static bool same_op_mod1(const EngineOp& a, const EngineOp& b)
{
    for (int i = 0; i < 9; i++)
        if (a.rf[i] != b.rf[i]) return false;
    for (int i = 0; i < 3; i++) {
        const double d = a.tf[i] - b.tf[i];
        if (std::fabs(d - std::nearbyint(d)) > 1e-9) return false;
    }
    return true;
}

// Coset decomposition of the explicit operations over their diagonal subgroup H: every operation
// is s = h o g for exactly one h in H and one representative g (the identity first), so
// sum_s X(s(x)) = sum_g sum_h X(h(g(x))) and the calculation can run over the images g(x) of the
// atoms with the operations of H, on the parity-folded kernels of H's sign set.  Built only when H
// has two or four operations without the inversion, some operation lies outside H and no
// operation conjugates the anomalous term.
static bool coset_plan(const std::vector<EngineOp>& ops, std::vector<EngineOp>* h_ops, std::vector<EngineOp>* reps)
{
    h_ops->clear();
    reps->clear();
    for (const EngineOp& o : ops) {
        if (o.anom_sign != 1.0) return false;
        const int p = sign_pattern(o);
        if (p == 7) return false;
        if (p >= 0) h_ops->push_back(o);
    }
    const size_t nh = h_ops->size();
    if ((nh != 2 && nh != 4) || ops.size() <= nh || ops.size() % nh != 0) return false;
    for (size_t a = 0; a < nh; a++)
        for (size_t b = a + 1; b < nh; b++)
            if (sign_pattern((*h_ops)[a]) == sign_pattern((*h_ops)[b])) return false;
    EngineOp id;
    for (int i = 0; i < 9; i++) id.rf[i] = i % 4 == 0 ? 1 : 0, id.rc[i] = i % 4 == 0 ? 1.0 : 0.0;
    for (int i = 0; i < 3; i++) id.tf[i] = 0.0;
    id.anom_sign = 1.0;
    std::vector<bool> covered(ops.size(), false);
    auto cover = [&](const EngineOp& g) {
        for (const EngineOp& h : *h_ops) {
            const EngineOp c = compose_ops(h, g);
            size_t j = 0;
            while (j < ops.size() && !same_op_mod1(c, ops[j])) j++;
            if (j == ops.size() || covered[j]) return false;
            covered[j] = true;
        }
        reps->push_back(g);
        return true;
    };
    if (!cover(id)) return false;
    for (size_t s = 0; s < ops.size(); s++)
        if (!covered[s] && !cover(ops[s])) return false;
    return reps->size() * nh == ops.size();
}

// The images g(x) of every atom under the coset representatives, image k of atom i at i K + k:
// position R_g x + t_g, U -> R U R^T (Cartesian), local frame R L, the same type, weight and f', f''.
static void expand_atoms(const std::vector<EngineOp>& reps, const EngineAtoms& a, EngineAtoms* x)
{
    const int K = (int)reps.size(), n = a.n;
    x->n = n * K;
    x->xyz_frac.resize(3 * (size_t)x->n);
    x->u6.resize(6 * (size_t)x->n);
    x->iso.resize(x->n);
    x->weight.resize(x->n);
    x->anom.resize(2 * (size_t)x->n);
    x->lcs.resize(9 * (size_t)x->n);
    x->type.resize(x->n);
    x->include.resize(x->n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < K; k++) {
            const EngineOp& g = reps[k];
            const size_t o = (size_t)i * K + k;
            for (int r = 0; r < 3; r++) {
                double v = g.tf[r];
                for (int c = 0; c < 3; c++) v += g.rf[3 * r + c] * a.xyz_frac[3 * (size_t)i + c];
                x->xyz_frac[3 * o + r] = v;
            }
            const double* u = &a.u6[6 * (size_t)i];
            if (a.iso[i]) {
                for (int j = 0; j < 6; j++) x->u6[6 * o + j] = u[j];
            } else {
                const double U[9] = {u[0], u[3], u[4], u[3], u[1], u[5], u[4], u[5], u[2]};
                double RU[9], V[9];
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) {
                        double s = 0.0;
                        for (int m = 0; m < 3; m++) s += g.rc[3 * r + m] * U[3 * m + c];
                        RU[3 * r + c] = s;
                    }
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) {
                        double s = 0.0;
                        for (int m = 0; m < 3; m++) s += RU[3 * r + m] * g.rc[3 * c + m];
                        V[3 * r + c] = s;
                    }
                const double v6[6] = {V[0], V[4], V[8], 0.5 * (V[1] + V[3]), 0.5 * (V[2] + V[6]), 0.5 * (V[5] + V[7])};
                for (int j = 0; j < 6; j++) x->u6[6 * o + j] = v6[j];
            }
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    double s = 0.0;
                    for (int m = 0; m < 3; m++) s += g.rc[3 * r + m] * a.lcs[9 * (size_t)i + 3 * m + c];
                    x->lcs[9 * o + 3 * r + c] = s;
                }
            x->iso[o] = a.iso[i];
            x->weight[o] = a.weight[i];
            x->anom[2 * o] = a.anom[2 * (size_t)i];
            x->anom[2 * o + 1] = a.anom[2 * (size_t)i + 1];
            x->type[o] = a.type[i];
            x->include[o] = a.include[i];
        }
}

// Derivative rows of the images summed back onto their atoms: dT/dx = sum_k R_k^T dT/dx_k
// (fractional), dT/dU = sum_k R_k^T G_k R_k with G_k the symmetric gradient of the image's U
// (off-diagonal entries of the 6-vector count twice), occupancy and the multipole rows (local
// frame, Pval, kappa, kappa') added.
static void fold_rows(const std::vector<EngineOp>& reps, int n, int nd, const double* img, double* out)
{
    const int K = (int)reps.size();
    for (int i = 0; i < n; i++) {
        double* r = out + (size_t)nd * i;
        for (int j = 0; j < nd; j++) r[j] = 0.0;
        for (int k = 0; k < K; k++) {
            const EngineOp& g = reps[k];
            const double* q = img + (size_t)nd * ((size_t)i * K + k);
            for (int c = 0; c < 3; c++)
                for (int m = 0; m < 3; m++) r[c] += g.rf[3 * m + c] * q[m];
            const double G[9] = {q[3], 0.5 * q[6], 0.5 * q[7], 0.5 * q[6], q[4], 0.5 * q[8], 0.5 * q[7], 0.5 * q[8], q[5]};
            double GR[9], M[9];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    double s = 0.0;
                    for (int m = 0; m < 3; m++) s += G[3 * a + m] * g.rc[3 * m + b];
                    GR[3 * a + b] = s;
                }
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    double s = 0.0;
                    for (int m = 0; m < 3; m++) s += g.rc[3 * m + a] * GR[3 * m + b];
                    M[3 * a + b] = s;
                }
            r[3] += M[0];
            r[4] += M[4];
            r[5] += M[8];
            r[6] += M[1] + M[3];
            r[7] += M[2] + M[6];
            r[8] += M[5] + M[7];
            for (int j = 9; j < nd; j++) r[j] += q[j];
        }
    }
}

// Gradient of T with respect to each atom's frame matrix L from its images' (frame R_k L):
// dT/dL = sum_k R_k^T dT/dL_k.
static void fold_frames(const std::vector<EngineOp>& reps, int n, const double* img, double* out)
{
    const int K = (int)reps.size();
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double s = 0.0;
                for (int k = 0; k < K; k++)
                    for (int m = 0; m < 3; m++)
                        s += reps[k].rc[3 * m + a] * img[9 * ((size_t)i * K + k) + 3 * m + b];
                out[9 * (size_t)i + 3 * a + b] = s;
            }
}

static std::vector<EngineOp> make_ops(const Cell& c, const SpaceGroupInfo& g, bool expand_centre,
                                      double partner_anom = -1.0)
{
    std::vector<EngineOp> ops;
    auto cart = [&](const int* rf, double* rc) {
        double t[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double s = 0;
                for (int k = 0; k < 3; k++) s += c.f2c[3 * i + k] * rf[3 * k + j];
                t[3 * i + j] = s;
            }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double s = 0;
                for (int k = 0; k < 3; k++) s += t[3 * i + k] * c.c2f[3 * k + j];
                rc[3 * i + j] = s;
            }
    };
    for (const SymOp& o : g.subset) {
        EngineOp e;
        std::memcpy(e.rf, o.r, sizeof(e.rf));
        std::memcpy(e.tf, o.t, sizeof(e.tf));
        cart(e.rf, e.rc);
        e.anom_sign = 1.0;
        ops.push_back(e);
    }
    if (expand_centre) {
        // inversion partners (-R, -t + t_inv); Engine3's term2 conjugates the whole atomic form
        // factor (HansenCoppens_SF_Engine3.cpp:2811-2813), hence anom_sign -1 there
        size_t n = ops.size();
        for (size_t s = 0; s < n; s++) {
            EngineOp e = ops[s];
            for (int i = 0; i < 9; i++) {
                e.rf[i] = -e.rf[i];
                e.rc[i] = -e.rc[i];
            }
            for (int i = 0; i < 3; i++) e.tf[i] = -ops[s].tf[i] + g.inv_t[i];
            e.anom_sign = partner_anom;
            ops.push_back(e);
        }
    }
    return ops;
}

static Calculator* create(const dh_model_desc* d, int device)
{
    if (!d) throw Error(DH_E_ARGUMENT, "null model description");
    std::unique_ptr<Calculator> c(new Calculator);
    c->cell.set(d->cell);
    if (!(d->cell[0] > 0 && d->cell[1] > 0 && d->cell[2] > 0)) throw Error(DH_E_MODEL, "invalid unit cell");
    std::vector<SymOp> ops(d->n_symops);
    for (int s = 0; s < d->n_symops; s++) {
        std::memcpy(ops[s].r, d->symop_rot + 9 * s, 9 * sizeof(int));
        std::memcpy(ops[s].t, d->symop_trans + 3 * s, 3 * sizeof(double));
    }
    c->sg = analyse_space_group(ops);
    c->xyz_conv = d->xyz_convention;
    c->adp_conv = d->adp_convention;
    c->d_xyz_conv = d->d_xyz_convention;
    c->d_adp_conv = d->d_adp_convention;
    // wavefunction types
    std::vector<WfnType> wfn(d->n_wfn);
    for (int w = 0; w < d->n_wfn; w++) {
        WfnType& t = wfn[w];
        int o0 = d->wfn_orbital_offset[w], o1 = d->wfn_orbital_offset[w + 1];
        for (int o = o0; o < o1; o++) {
            SlaterOrbital so;
            so.n = d->orbital_nl[2 * o];
            so.l = d->orbital_nl[2 * o + 1];
            for (int k = d->orbital_term_offset[o]; k < d->orbital_term_offset[o + 1]; k++) {
                so.coeff.push_back(d->term_coeff[k]);
                so.expo.push_back(d->term_exp[k]);
                so.pow.push_back(d->term_pow[k]);
            }
            int local = o - o0;
            t.orbitals.push_back(so);
            t.occupancy.push_back(d->orbital_occupancy[o]);
            if (d->orbital_kind[o] == 0) t.core.push_back(local);
            if (d->orbital_kind[o] == 1) t.valence.push_back(local);
        }
        t.def_val_exponent = d->def_val_exponent[w];
        for (int k = d->def_val_pow_offset[w]; k < d->def_val_pow_offset[w + 1]; k++)
            t.def_val_power.push_back(d->def_val_pow[k]);
        t.anomalous = {d->anomalous ? d->anomalous[2 * w] : 0.0, d->anomalous ? d->anomalous[2 * w + 1] : 0.0};
    }
    std::vector<AtomType> types(d->n_types);
    for (int t = 0; t < d->n_types; t++) {
        types[t].n_l = d->plm_n_l[t];
        for (int i = 0; i < 25; i++) types[t].plm[i] = d->plm[25 * t + i];
        types[t].p_val = d->p_val[t];
        types[t].kappa = d->kappa[t];
        types[t].kappa_prime = d->kappa_prime[t];
        if (!(types[t].kappa > 0) || !(types[t].kappa_prime > 0)) throw Error(DH_E_MODEL, "kappa must be positive");
    }
    c->atom_type.assign(d->atom_to_type, d->atom_to_type + d->n_atoms);
    c->atom_wfn.assign(d->atom_to_wfn, d->atom_to_wfn + d->n_atoms);
    c->model = build_engine_model(wfn, types, c->atom_wfn, c->atom_type, &c->eng_atom_type);
    c->types = types;
    c->mult_w.resize(d->n_atoms);
    for (int a = 0; a < d->n_atoms; a++) c->mult_w[a] = d->atom_multiplicity[a] / double(c->sg.order);
    bool zero_t = c->sg.inv_t[0] == 0.0 && c->sg.inv_t[1] == 0.0 && c->sg.inv_t[2] == 0.0;
    c->centric_origin = c->sg.centro && zero_t;
    bool centric_shift = c->sg.centro && !zero_t;
    std::vector<EngineOp> eops = make_ops(c->cell, c->sg, centric_shift);
    c->engine.reset(new Engine(device, c->model, d->cell, eops, c->centric_origin));
    std::vector<EngineOp> h_ops;
    if (coset_plan(eops, &h_ops, &c->cos_reps))
        c->cos_engine.reset(new Engine(device, c->model, d->cell, h_ops, c->centric_origin));
    else
        c->cos_reps.clear();
    c->device = device;
    for (int w = 0; w < d->n_wfn; w++) {
        c->wfn_label.push_back(d->wfn_label && d->wfn_label[w] ? std::string(d->wfn_label[w]) : std::string());
        c->wfn_z.push_back(d->wfn_atomic_number ? d->wfn_atomic_number[w] : 0);
        c->wfn_anom.push_back(wfn[w].anomalous);
    }
    return c.release();
}

static void ensure_iam_engine(Calculator* c)
{
    if (c->iam_engine) return;
    for (const std::string& l : c->wfn_label)
        if (l.empty()) throw Error(DH_E_MODEL, "CPU_IAM needs the wavefunction type labels (element symbols)");
    EngineModel m = build_iam_model(c->wfn_label, c->wfn_anom, c->atom_wfn, c->iam_atom_type, c->iam_table);
    const bool centro = c->sg.centro;
    std::vector<EngineOp> ops = make_ops(c->cell, c->sg, centro, +1.0);
    c->iam_engine.reset(new Engine(c->device, m, c->cell.par, ops, false));
}

// rows: inputs of the per-reflection derivative rows (no dT/dF; reflections kept in caller order)
static void upload(Calculator* c, const dh_atoms* at, int n_hkl, const int* hkl, const double* dtdf,
                   bool rows = false)
{
    // whatever fails below, the engines keep no inputs of an earlier call to run on
    c->invalidate_all();
    Engine& E = rows ? c->eng() : c->host_eng();
    if (!at) throw Error(DH_E_ARGUMENT, "null atoms");
    const int n = at->n_atoms;
    if (n != (int)c->atom_type.size())
        throw Error(DH_E_ARGUMENT, "number of atoms differs from the model's atom-to-type map");
    if (n_hkl < 0) throw Error(DH_E_ARGUMENT, "negative number of reflections");
    if (n_hkl > 0 && !hkl) throw Error(DH_E_ARGUMENT, "null Miller indices");
    if (n_hkl > 0 && !dtdf && !rows)
        throw Error(DH_E_ARGUMENT,
                    "Inconsistent size of hkl vectors set and (d target/ d F) table when calculating multipolar "
                    "structure factors.");
    c->rows_uploaded = false;
    const double two_pi2 = 2.0 * M_PI * M_PI;
    // kept in the calculator between calls: every entry is written below, so a call with the
    // sizes of the last one allocates and zeroes nothing
    EngineAtoms& ea = c->up_atoms;
    ea.n = n;
    ea.xyz_frac.resize(3 * n);
    ea.u6.resize(6 * n);
    ea.iso.resize(n);
    ea.weight.resize(n);
    ea.anom.resize(2 * n);
    ea.lcs.resize(9 * n);
    ea.type = c->iam ? c->iam_atom_type : c->eng_atom_type;
    ea.include.resize(n);
    c->adp_size.resize(n);
    c->occ.resize(n);
    for (int i = 0; i < n; i++) {
        const double* x = at->xyz + 3 * i;
        if (c->xyz_conv == DH_XYZ_CARTESIAN) {
            for (int j = 0; j < 3; j++)
                ea.xyz_frac[3 * i + j] = c->cell.c2f[3 * j] * x[0] + c->cell.c2f[3 * j + 1] * x[1] + c->cell.c2f[3 * j + 2] * x[2];
        } else {
            for (int j = 0; j < 3; j++) ea.xyz_frac[3 * i + j] = x[j];
        }
        int na = at->adp_size[i];
        if (na != 0 && na != 1 && na != 6) throw Error(DH_E_ARGUMENT, "ADP must have 0, 1 or 6 components");
        c->adp_size[i] = na;
        const double* u = at->adp + 6 * i;
        ea.iso[i] = 0;
        for (int j = 0; j < 6; j++) ea.u6[6 * i + j] = 0.0;
        if (na == 6) {
            double uc[6];
            if (c->adp_conv == DH_ADP_U_CART) std::memcpy(uc, u, sizeof(uc));
            else c->cell.u_cif_to_cart(u, uc);
            for (int j = 0; j < 6; j++) ea.u6[6 * i + j] = uc[j] * two_pi2;
        } else if (na == 1) {
            ea.iso[i] = 1;
            for (int j = 0; j < 3; j++) ea.u6[6 * i + j] = u[0] * two_pi2;
        }
        c->occ[i] = at->occupancy[i];
        ea.weight[i] = at->occupancy[i] * c->mult_w[i];
        int w = c->atom_wfn[i];
        if (at->anomalous) {
            ea.anom[2 * i] = at->anomalous[2 * i];
            ea.anom[2 * i + 1] = at->anomalous[2 * i + 1];
        } else {
            ea.anom[2 * i] = c->model.anom[2 * w];
            ea.anom[2 * i + 1] = c->model.anom[2 * w + 1];
        }
        for (int j = 0; j < 9; j++) ea.lcs[9 * i + j] = at->lcs ? at->lcs[9 * i + j] : (j % 4 == 0 ? 1.0 : 0.0);
        ea.include[i] = at->include ? at->include[i] : 1;
    }
    if (!c->lcs_defs.empty()) {
        // AnyHcCalculator::update (AnyHcCalculator.cpp:317-330): frames from the current positions
        parallel_atoms(n, 4096, [&](int a0, int a1) {
            for (int i = a0; i < a1; i++)
                lcs_matrix(c->lcs_defs[i], ea.xyz_frac.data(), c->cell.f2c, &ea.lcs[9 * (size_t)i]);
        });
    }
    c->last_lcs = ea.lcs;
    c->last_xyz_frac = ea.xyz_frac;
    // centering multipliers, dT/dF <- conj(dT/dF) * mult  (.cpp:393-404); the multipliers depend
    // on the reflections only and are kept for a list equal to the last one
    if (c->mult_hkl.size() != 3 * (size_t)n_hkl ||
        (n_hkl > 0 && std::memcmp(c->mult_hkl.data(), hkl, 3 * (size_t)n_hkl * sizeof(int)) != 0)) {
        c->mult_hkl.assign(hkl, hkl + 3 * (size_t)n_hkl);
        c->mult_k.resize(n_hkl);
        for (int k = 0; k < n_hkl; k++) c->mult_k[k] = centering_multiplier(c->sg.centering, c->sg.obverse, hkl + 3 * k);
    }
    std::vector<double>& dd = c->up_dtdf;
    dd.resize(3 * (size_t)n_hkl);
    // F = 2 Re(sum) and derivatives 2 Re(c) Re(.) for a centre at the origin (Engine3.cpp:1924-1958)
    const bool centric = c->eng_centric_origin();
    const double* mk = c->mult_k.data();
    if (rows) {
        for (int k = 0; k < 2 * n_hkl; k++) dd[k] = 0.0;
    } else if (centric) {
        for (int k = 0; k < n_hkl; k++) {
            dd[2 * k] = dtdf[2 * k] * mk[k] * 2.0;
            dd[2 * k + 1] = 0.0;
        }
    } else {
        for (int k = 0; k < n_hkl; k++) {
            dd[2 * k] = dtdf[2 * k] * mk[k];
            dd[2 * k + 1] = -dtdf[2 * k + 1] * mk[k];
        }
    }
    std::memcpy(dd.data() + 2 * (size_t)n_hkl, mk, (size_t)n_hkl * sizeof(double));
    E.line_order = true;     // both precisions carry phases along hkl lines (groups of <= 4 operations)
    if (!rows && c->cosets()) {
        expand_atoms(c->cos_reps, ea, &c->cos_atoms);
        E.upload(c->cos_atoms, n_hkl, hkl, dd.data(), rows);
    } else {
        E.upload(ea, n_hkl, hkl, dd.data(), rows);
    }
    c->n_atoms = n;
    c->n_hkl = n_hkl;
    if (!rows) return;
    // derivative conventions of the rows (WrapperForSingleHklResultProcessor::convert,
    // HansenCoppensStructureFactorCalculator.h:512-531): d/dx_cart = c2f^T d/dx_frac and the
    // U_cart -> U_cif chain rule, both linear, applied inside the kernel through its records
    JacConv jc;
    for (int i = 0; i < 3; i++)
        for (int m = 0; m < 3; m++)
            jc.xyz_m[3 * i + m] = c->d_xyz_conv == DH_XYZ_CARTESIAN ? c->cell.c2f[3 * m + i] : (i == m ? 1.0 : 0.0);
    for (int j = 0; j < 6; j++) {
        double e[6] = {0, 0, 0, 0, 0, 0}, o[6];
        e[j] = 1.0;
        if (c->d_adp_conv == DH_ADP_U_CART) std::memcpy(o, e, sizeof(o));
        else c->cell.d_ucart_to_d_ucif(e, o);
        for (int i = 0; i < 6; i++) jc.adp_m[6 * i + j] = o[i];
    }
    jc.mult_w = c->mult_w;
    jc.adp_kind = c->adp_size;
    c->eng().set_jacobian_conventions(jc);
    c->rows_uploaded = true;
}

static constexpr size_t kRowBlockBytes = 64u << 20;   // device / pinned staging per block

static int rows_per_block(const Calculator* c)
{
    const size_t row = ((size_t)DH_JAC_PLANES * c->n_atoms * 2 + 2) * sizeof(double);
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(c->n_hkl, 1), kRowBlockBytes / row));
}

static void check_rows(Calculator* c, int k0, int k1)
{
    if (!c->rows_uploaded) throw Error(DH_E_ARGUMENT, "per-reflection rows: call dh_hc_jacobian_upload first");
    if (k0 < 0 || k1 < k0 || k1 > c->n_hkl) throw Error(DH_E_ARGUMENT, "reflection range outside the upload");
}

// Rows [k0, k1) into host memory, through the calculator's device blocks.
static void rows_to_host(Calculator* c, int k0, int k1, double* f, double* rows)
{
    check_rows(c, k0, k1);
    const int B = rows_per_block(c);
    const size_t row = (size_t)DH_JAC_PLANES * c->n_atoms * 2;
    hipStream_t st = c->eng().stream();
    for (int b0 = k0; b0 < k1; b0 += B) {
        const int b1 = std::min(k1, b0 + B), n = b1 - b0;
        c->rows_dev[0].ensure((size_t)n * row * sizeof(double));
        c->f_dev[0].ensure((size_t)n * 2 * sizeof(double));
        c->eng().jacobian(b0, b1, c->f_dev[0].as<double>(), c->rows_dev[0].as<double>(), c->prec);
        c->last_run = &c->eng();
        if (f)
            HIP_CHECK(hipMemcpyAsync(f + 2 * (size_t)(b0 - k0), c->f_dev[0].p, (size_t)n * 2 * sizeof(double),
                                     hipMemcpyDeviceToHost, st));
        if (rows)
            HIP_CHECK(hipMemcpyAsync(rows + row * (b0 - k0), c->rows_dev[0].p, (size_t)n * row * sizeof(double),
                                     hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
}

// The per-hkl driver: blocks of rows computed on the GPU while the host hands the previous block
// to the callback, reflection by reflection in caller order.
static void per_hkl(Calculator* c, dh_per_hkl_fn fn, void* user)
{
    if (!fn) throw Error(DH_E_ARGUMENT, "null per-reflection callback");
    const int n_hkl = c->n_hkl;
    if (n_hkl == 0) return;
    const int B = rows_per_block(c);
    const size_t row = (size_t)DH_JAC_PLANES * c->n_atoms * 2;
    hipStream_t st = c->eng().stream();
    for (int i = 0; i < 2; i++) {
        if (!c->rows_ev[i]) HIP_CHECK(hipEventCreateWithFlags(&c->rows_ev[i], hipEventDisableTiming));
        c->rows_dev[i].ensure((size_t)B * row * sizeof(double));
        c->f_dev[i].ensure((size_t)B * 2 * sizeof(double));
        c->rows_host[i].ensure((size_t)B * row * sizeof(double));
        c->f_host[i].ensure((size_t)B * 2 * sizeof(double));
    }
    const int nb = (n_hkl + B - 1) / B;
    for (int b = 0; b <= nb; b++) {
        if (b < nb) {
            const int q = b & 1, b0 = b * B, b1 = std::min(n_hkl, b0 + B), n = b1 - b0;
            c->eng().jacobian(b0, b1, c->f_dev[q].as<double>(), c->rows_dev[q].as<double>(), c->prec);
            c->last_run = &c->eng();
            HIP_CHECK(hipMemcpyAsync(c->f_host[q].p, c->f_dev[q].p, (size_t)n * 2 * sizeof(double),
                                     hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipMemcpyAsync(c->rows_host[q].p, c->rows_dev[q].p, (size_t)n * row * sizeof(double),
                                     hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipEventRecord(c->rows_ev[q], st));
        }
        if (b > 0) {
            const int q = (b - 1) & 1, b0 = (b - 1) * B, b1 = std::min(n_hkl, b0 + B);
            HIP_CHECK(hipEventSynchronize(c->rows_ev[q]));
            for (int k = b0; k < b1; k++) {
                const int rc = fn(user, k, c->f_host[q].d() + 2 * (size_t)(k - b0), c->rows_host[q].d() + row * (k - b0));
                if (rc != 0) {
                    HIP_CHECK(hipStreamSynchronize(st));
                    throw Error(DH_E_CALLBACK, "per-reflection callback returned " + std::to_string(rc));
                }
            }
        }
    }
}

static void fetch(Calculator* c, double* f, double* d_xyz, double* d_adp, double* d_occ)
{
    const int n = c->n_atoms;
    const bool want_d = c->last_selector != 0;
    Engine& E = c->host_eng();
    const int nd = E.nd();
    c->d10.assign((size_t)std::max(n, 1) * nd, 0.0);
    if (c->cosets() && want_d) {
        // the images' rows, folded back onto their atoms
        c->cos_rows.resize((size_t)std::max(n, 1) * c->cos_reps.size() * nd);
        E.download(f, c->cos_rows.data());
        fold_rows(c->cos_reps, n, nd, c->cos_rows.data(), c->d10.data());
    } else {
        E.download(f, want_d ? c->d10.data() : nullptr);
    }
    c->have_mp = false;
    if (!want_d) {
        if (d_xyz) std::memset(d_xyz, 0, sizeof(double) * 3 * n);
        if (d_adp) std::memset(d_adp, 0, sizeof(double) * 6 * n);
        if (d_occ) std::memset(d_occ, 0, sizeof(double) * n);
        return;
    }
    const double two_pi = 2.0 * M_PI, two_pi2 = 2.0 * M_PI * M_PI;
    // chain rule through the recomputed frames: dT/dx_frac += w_a dT/dL_a : dL_a/dx_frac
    std::vector<double> chain;
    if (c->lcs_chain && !c->lcs_defs.empty() && (c->last_selector & DH_D_XYZ)) {
        std::vector<double> g(9 * (size_t)n);
        if (c->cosets()) {
            std::vector<double> gi(9 * (size_t)n * c->cos_reps.size());
            E.download_lcs_grad(gi.data());
            fold_frames(c->cos_reps, n, gi.data(), g.data());
        } else {
            E.download_lcs_grad(g.data());
        }
        chain.assign(3 * (size_t)n, 0.0);
        // per-frame contributions on host threads, then one serial scatter
        std::vector<int> who((size_t)n * kLcsMaxAtoms), cnt(n);
        std::vector<double> part((size_t)n * kLcsMaxAtoms * 3);
        auto work = [&](int a0, int a1) {
            for (int a = a0; a < a1; a++) {
                double gw[9];
                const double w = c->occ[a] * c->mult_w[a];
                for (int k = 0; k < 9; k++) gw[k] = w * g[9 * (size_t)a + k];
                cnt[a] = lcs_chain(c->lcs_defs[a], c->last_xyz_frac.data(), c->cell.f2c, gw,
                                   &who[(size_t)a * kLcsMaxAtoms], &part[(size_t)a * kLcsMaxAtoms * 3]);
            }
        };
        parallel_atoms(n, 512, work);
        for (int a = 0; a < n; a++)
            for (int s = 0; s < cnt[a]; s++)
                for (int j = 0; j < 3; j++)
                    chain[3 * (size_t)who[(size_t)a * kLcsMaxAtoms + s] + j] += part[((size_t)a * kLcsMaxAtoms + s) * 3 + j];
    }
    for (int i = 0; i < n; i++) {
        const double* r = c->d10.data() + (size_t)nd * i;
        double w = c->occ[i] * c->mult_w[i];
        double dxf[3] = {two_pi * w * r[0], two_pi * w * r[1], two_pi * w * r[2]};
        if (!chain.empty())
            for (int j = 0; j < 3; j++) dxf[j] += chain[3 * (size_t)i + j];
        if (d_xyz) {
            if (c->d_xyz_conv == DH_XYZ_CARTESIAN) {
                // d/dx_cart = c2f^T d/dx_frac
                for (int j = 0; j < 3; j++)
                    d_xyz[3 * i + j] = c->cell.c2f[j] * dxf[0] + c->cell.c2f[3 + j] * dxf[1] + c->cell.c2f[6 + j] * dxf[2];
            } else {
                for (int j = 0; j < 3; j++) d_xyz[3 * i + j] = dxf[j];
            }
            if (!(c->last_selector & DH_D_XYZ)) d_xyz[3 * i] = d_xyz[3 * i + 1] = d_xyz[3 * i + 2] = 0.0;
        }
        if (d_adp) {
            double* o = d_adp + 6 * (size_t)i;
            for (int j = 0; j < 6; j++) o[j] = 0.0;
            if (c->last_selector & DH_D_ADP) {
                double du[6];
                for (int j = 0; j < 6; j++) du[j] = two_pi2 * w * r[3 + j];
                if (c->adp_size[i] == 1) o[0] = du[0] + du[1] + du[2];
                else if (c->adp_size[i] == 6) {
                    if (c->d_adp_conv == DH_ADP_U_CART) std::memcpy(o, du, sizeof(du));
                    else c->cell.d_ucart_to_d_ucif(du, o);
                }
            }
        }
        if (d_occ) d_occ[i] = (c->last_selector & DH_D_OCC) ? c->mult_w[i] * r[9] : 0.0;
    }
    if ((c->last_selector & DH_D_MULTIPOLE) && nd >= 38) {
        const int nt = (int)c->types.size();
        c->d_plm.assign(25 * (size_t)nt, 0.0);
        c->d_pval.assign(nt, 0.0);
        c->d_kappa.assign(nt, 0.0);
        c->d_kappa_prime.assign(nt, 0.0);
        for (int i = 0; i < n; i++) {
            const double* r = c->d10.data() + (size_t)nd * i;
            const double w = c->occ[i] * c->mult_w[i];
            const int t = c->atom_type[i];
            for (int j = 0; j < 25; j++) c->d_plm[25 * t + j] += w * r[10 + j];
            c->d_pval[t] += w * r[35];
            c->d_kappa[t] += w * r[36];
            c->d_kappa_prime[t] += w * r[37];
        }
        // only entries present in the type's p_lm array are parameters
        for (int t = 0; t < nt; t++)
            for (int l = 0; l < 5; l++)
                if (l >= c->types[t].n_l)
                    for (int j = l * l; j < (l + 1) * (l + 1); j++) c->d_plm[25 * t + j] = 0.0;
        c->have_mp = true;
    }
}

// Device-resident calls: the layout upload (every atom in its slot, reflections from the host)
// and the conversion matrices between the calculator's conventions and the engine's.
static void set_reflections(Calculator* c, int n_hkl, const int* hkl)
{
    c->invalidate_all();
    Engine& E = c->host_eng();         // the engine of the summed calls (its coset engine where it has one)
    c->rows_uploaded = false;
    if (n_hkl < 0) throw Error(DH_E_ARGUMENT, "negative number of reflections");
    if (n_hkl > 0 && !hkl) throw Error(DH_E_ARGUMENT, "null Miller indices");
    const int n = (int)c->atom_type.size();
    EngineAtoms ea;
    ea.n = n;
    ea.xyz_frac.assign(3 * (size_t)n, 0.0);
    ea.u6.assign(6 * (size_t)n, 0.0);
    ea.iso.assign(n, 0);
    ea.weight.assign(n, 0.0);
    ea.anom.assign(2 * (size_t)n, 0.0);
    ea.lcs.assign(9 * (size_t)n, 0.0);
    for (int i = 0; i < n; i++) ea.lcs[9 * (size_t)i] = ea.lcs[9 * (size_t)i + 4] = ea.lcs[9 * (size_t)i + 8] = 1.0;
    ea.type = c->iam ? c->iam_atom_type : c->eng_atom_type;
    ea.include.assign(n, 1);
    std::vector<double> mult(n_hkl), anom(2 * (size_t)n);
    for (int k = 0; k < n_hkl; k++) mult[k] = centering_multiplier(c->sg.centering, c->sg.obverse, hkl + 3 * k);
    for (int i = 0; i < n; i++) {
        anom[2 * (size_t)i] = c->model.anom[2 * c->atom_wfn[i]];
        anom[2 * (size_t)i + 1] = c->model.anom[2 * c->atom_wfn[i] + 1];
    }
    const double two_pi2 = 2.0 * M_PI * M_PI;
    DevConv conv;
    conv.xyz_identity = c->xyz_conv != DH_XYZ_CARTESIAN;
    std::memcpy(conv.c2f, c->cell.c2f, sizeof(conv.c2f));
    conv.dxyz_identity = c->d_xyz_conv != DH_XYZ_CARTESIAN;
    for (int j = 0; j < 3; j++)
        for (int m = 0; m < 3; m++) conv.dxyz_m[3 * j + m] = c->cell.c2f[3 * m + j];   // d/dx_cart = c2f^T d/dx_frac
    for (int k = 0; k < 6; k++) {
        double e[6] = {0, 0, 0, 0, 0, 0}, uc[6], o[6];
        e[k] = 1.0;
        if (c->adp_conv == DH_ADP_U_CART) std::memcpy(uc, e, sizeof(uc));
        else c->cell.u_cif_to_cart(e, uc);
        if (c->d_adp_conv == DH_ADP_U_CART) std::memcpy(o, e, sizeof(o));
        else c->cell.d_ucart_to_d_ucif(e, o);
        for (int j = 0; j < 6; j++) {
            conv.u2c[6 * j + k] = uc[j] * two_pi2;
            conv.dadp_m[6 * j + k] = o[j];
        }
    }
    conv.centric_origin = c->eng_centric_origin() ? 1 : 0;
    E.line_order = true;
    std::vector<int> sum_type = c->atom_type;
    if (c->cosets()) {
        // K images of every atom (image k of atom i at i K + k), built per call from the caller's
        // arrays by k_dev_atoms
        const int K = (int)c->cos_reps.size();
        EngineAtoms im;
        expand_atoms(c->cos_reps, ea, &im);
        std::vector<double> mw((size_t)n * K), an(2 * (size_t)n * K);
        sum_type.resize((size_t)n * K);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < K; k++) {
                const size_t o = (size_t)i * K + k;
                mw[o] = c->mult_w[i];
                an[2 * o] = anom[2 * (size_t)i];
                an[2 * o + 1] = anom[2 * (size_t)i + 1];
                sum_type[o] = c->atom_type[i];
            }
        E.prepare_dev(im, n_hkl, hkl, mult.data(), mw, an, conv, c->prec);
        E.set_dev_cosets(c->cos_reps);
    } else {
        E.prepare_dev(ea, n_hkl, hkl, mult.data(), c->mult_w, anom, conv, c->prec);
        E.set_dev_cosets({});
    }
    if (!c->iam) {
        // the per-type sums of the multipole derivatives (dh_hc_sf_and_derivs_mp_dev)
        std::vector<int> n_l(c->types.size());
        for (size_t t = 0; t < c->types.size(); t++) n_l[t] = c->types[t].n_l;
        E.set_type_sum_map(sum_type, n_l);
    }
    c->n_atoms = n;
    c->n_hkl = n_hkl;
}

static void check_selector(Calculator* c, unsigned selector);

// mp: the multipole-parameter derivatives as well (dh_hc_sf_and_derivs_mp_dev), summed per type
// into the four device arrays of `tsum`
static void sf_dev(Calculator* c, const dh_atoms_dev* at, const double* dtdf, unsigned selector, double* f,
                   double* d_xyz, double* d_adp, double* d_occ, void* stream, bool no_alloc, bool mp = false,
                   double* const* tsum = nullptr)
{
    if (!at) throw Error(DH_E_ARGUMENT, "null atoms");
    if (at->n_atoms != (int)c->atom_type.size())
        throw Error(DH_E_ARGUMENT, "number of atoms differs from the model's atom-to-type map");
    if (mp) {
        if (selector & ~(DH_D_ALL | DH_D_MULTIPOLE))
            throw Error(DH_E_ARGUMENT, "unknown derivative selector bits");
        check_selector(c, selector);
        if (!(selector & DH_D_MULTIPOLE)) mp = false;
        selector &= DH_D_ALL;
    } else if (selector & ~DH_D_ALL) {
        throw Error(DH_E_ARGUMENT, "dh_hc_sf_and_derivs_dev takes DH_D_XYZ, DH_D_ADP and DH_D_OCC (multipole "
                                   "derivatives: dh_hc_sf_and_derivs_mp_dev or the host call)");
    }
    if (!c->lcs_defs.empty())
        throw Error(DH_E_ARGUMENT, "device calls take the local frames as given (frames recomputed from positions "
                                   "need the host call)");
    DevCall call;
    call.xyz = at->xyz;
    call.adp_size = at->adp_size;
    call.adp = at->adp;
    call.occ = at->occupancy;
    call.lcs = at->lcs;
    call.anom = at->anomalous;
    call.include = at->include;
    call.dtdf = dtdf;
    call.f = f;
    call.d_xyz = d_xyz;
    call.d_adp = d_adp;
    call.d_occ = d_occ;
    call.selector = selector;
    if (mp) {
        call.multipole = true;
        call.d_plm = tsum[0];
        call.d_pval = tsum[1];
        call.d_kappa = tsum[2];
        call.d_kappa_prime = tsum[3];
    }
    Engine& e = c->host_eng();
    e.run_dev(call, c->prec, static_cast<hipStream_t>(stream), no_alloc ? e.scratch() : nullptr,
              no_alloc ? Engine::scratch_bytes() : 0);
}

// the multipole-adjoint kernels run when asked for, and for the chain rule through the frames
static bool run_mp(Calculator* c, unsigned selector)
{
    const bool chain = c->lcs_chain && !c->lcs_defs.empty() && (selector & DH_D_XYZ) && !c->iam;
    c->host_eng().lcs_grad = chain;
    c->host_eng().mp_scalars = (selector & DH_D_MULTIPOLE) != 0;
    return (selector & DH_D_MULTIPOLE) != 0 || chain;
}

static LcsRef to_ref(const dh_atom_ref& r)
{
    LcsRef o;
    o.atom = r.atom;
    for (int i = 0; i < 9; i++) o.rot[i] = r.rot[i];
    for (int i = 0; i < 3; i++) o.trans[i] = r.trans[i];
    return o;
}

static std::vector<LcsDef> to_defs(int n, const dh_lcs_def* defs, int n_atoms)
{
    if (n < 0 || (n > 0 && !defs)) throw Error(DH_E_ARGUMENT, "invalid local coordinate system definitions");
    std::vector<LcsDef> out(n);
    for (int i = 0; i < n; i++) {
        const dh_lcs_def& d = defs[i];
        LcsDef& o = out[i];
        o.central = to_ref(d.central);
        o.dir1 = d.direction1_type;
        o.dir2 = d.direction2_type;
        if (d.n_ref1 < 0 || d.n_ref2 < 0 || (d.n_ref1 > 0 && !d.ref1) || (d.n_ref2 > 0 && !d.ref2) ||
            (d.n_chirality > 0 && !d.chirality))
            throw Error(DH_E_ARGUMENT, "invalid local coordinate system definitions");
        for (int k = 0; k < d.n_ref1; k++) o.ref1.push_back(to_ref(d.ref1[k]));
        for (int k = 0; k < d.n_ref2; k++) o.ref2.push_back(to_ref(d.ref2[k]));
        for (int k = 0; k < d.n_chirality; k++) o.chirality.push_back(to_ref(d.chirality[k]));
        o.c1 = d.coordinate_1;
        o.c2 = d.coordinate_2;
        o.is_r = d.is_r != 0;
        lcs_validate(o, n_atoms);
    }
    return out;
}

static void check_selector(Calculator* c, unsigned selector)
{
    if (c->iam && (selector & DH_D_MULTIPOLE))
        throw Error(DH_E_ARGUMENT, "multipole-parameter derivatives need a multipole model (not CPU_IAM)");
}

// Electron scattering per atom type: Z of the type's wavefunction and the electron form factor at
// zero angle from the "electron-IT" table (HansenCoppens_SF_Engine3::electronScatteringAt000,
// HansenCoppens_SF_Engine3.cpp:952-970).
static void set_electron(Calculator* c, bool on)
{
    std::vector<double> z_f0;
    if (on) {
        for (int t = 0; t < c->model.n_types; t++) {
            const int w = c->model.wfn_origin[c->model.type_wfn[t]];
            const int z = w < (int)c->wfn_z.size() ? c->wfn_z[w] : 0;
            if (z <= 0) throw Error(DH_E_MODEL, "electron scattering needs the atomic number of every wavefunction type");
            int n = 0;
            const NGaussEntry* e = ngauss_lookup(element_symbol(z), "electron-IT", &n);
            if (!e) throw Error(DH_E_MODEL, "no electron form factor for element " + element_symbol(z));
            z_f0.push_back((double)z);
            z_f0.push_back(ngauss_value(*e, n, 0.0));
        }
    }
    c->engine->set_electron(on, z_f0);
    if (c->cos_engine) c->cos_engine->set_electron(on, z_f0);
    c->electron = on;
}

}  // namespace discamb_amd

using namespace discamb_amd;

struct dh_calculator {
    Calculator* c;
};

struct dh_batch {
    std::vector<Calculator*> calcs;
    std::vector<Engine*> engines;     // the calculators' host-call engines at creation
    std::unique_ptr<BatchEngine> engine;
};

template <typename F>
static int guarded(F&& f)
{
    try {
        f();
        return DH_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return DH_E_ARGUMENT;
    }
}

extern "C" {

const char* dh_last_error(void) { return g_last_error.c_str(); }

const char* dh_version(void) { return "discamb_amd 0.1 (gfx950)"; }

int dh_device_count(int* n)
{
    return guarded([&] {
        int k = 0;
        if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
        *n = k;
    });
}

int dh_hc_create(const dh_model_desc* model, int device, dh_calculator** out)
{
    return guarded([&] {
        if (!out) throw Error(DH_E_ARGUMENT, "null output pointer");
        *out = nullptr;
        Calculator* c = create(model, device);
        *out = new dh_calculator{c};
    });
}

void dh_hc_destroy(dh_calculator* calc)
{
    if (!calc) return;
    delete calc->c;
    delete calc;
}

int dh_hc_set_calculation_mode(dh_calculator* calc, const char* mode)
{
    return guarded([&] {
        if (!calc || !mode) throw Error(DH_E_ARGUMENT, "null argument");
        std::string m(mode);
        Calculator* c = calc->c;
        if (m == "HIP" || m == "GPU" || m == "CPU" || m == "CPU_v4") {
            c->prec = kFp64;
            c->iam = false;
        } else if (m == "HIP_MIXED") {
            c->prec = kMixed;
            c->iam = false;
        } else if (m == "CPU_IAM" || m == "HIP_IAM" || m == "HIP_IAM_MIXED") {
            ensure_iam_engine(c);
            c->iam_engine->tune = c->engine->tune;
            c->iam_engine->timing = c->engine->timing;
            c->prec = m == "HIP_IAM_MIXED" ? kMixed : kFp64;
            c->iam = true;
        } else
            throw Error(DH_E_ARGUMENT, "unknown calculation mode for multipolar structure factor engine");
        c->rows_uploaded = false;
    });
}

int dh_hc_set_iam_table(dh_calculator* calc, const char* table)
{
    return guarded([&] {
        if (!calc || !table) throw Error(DH_E_ARGUMENT, "null argument");
        Calculator* c = calc->c;
        const std::string t(table);
        ngauss_lookup("H", t, nullptr);     // rejects an unknown table name
        if (t == c->iam_table) return;
        c->iam_table = t;
        if (c->last_run == c->iam_engine.get()) c->last_run = nullptr;
        c->iam_engine.reset();
        if (c->iam) {
            ensure_iam_engine(c);
            c->iam_engine->tune = c->engine->tune;
            c->iam_engine->timing = c->engine->timing;
        }
        c->rows_uploaded = false;
    });
}

int dh_lcs_compute(const double cell[6], int n_atoms, const double* xyz_frac, int n, const dh_lcs_def* defs, double* lcs)
{
    return guarded([&] {
        if (!cell || (n > 0 && (!xyz_frac || !lcs))) throw Error(DH_E_ARGUMENT, "null argument");
        Cell c;
        c.set(cell);
        std::vector<LcsDef> d = to_defs(n, defs, n_atoms);
        for (int i = 0; i < n; i++) lcs_matrix(d[i], xyz_frac, c.f2c, lcs + 9 * (size_t)i);
    });
}

int dh_hc_set_lcs_definitions(dh_calculator* calc, int n, const dh_lcs_def* defs, int chain_rule)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        Calculator* c = calc->c;
        if (n != 0 && n != (int)c->atom_type.size())
            throw Error(DH_E_ARGUMENT, "one local coordinate system definition per atom is needed");
        c->lcs_defs = to_defs(n, defs, (int)c->atom_type.size());
        c->lcs_chain = chain_rule != 0 && n > 0;
    });
}

int dh_hc_last_lcs(dh_calculator* calc, double* lcs)
{
    return guarded([&] {
        if (!calc || !lcs) throw Error(DH_E_ARGUMENT, "null argument");
        std::memcpy(lcs, calc->c->last_lcs.data(), calc->c->last_lcs.size() * sizeof(double));
    });
}

int dh_hc_set_electron_scattering(dh_calculator* calc, int on)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        set_electron(calc->c, on != 0);
    });
}

int dh_hc_set_tuning(dh_calculator* calc, int hkl_chunk)
{
    return guarded([&] {
        if (!calc || hkl_chunk < 0) throw Error(DH_E_ARGUMENT, "invalid tuning");
        calc->c->engine->tune.hkl_chunk = hkl_chunk;
        if (calc->c->iam_engine) calc->c->iam_engine->tune.hkl_chunk = hkl_chunk;
        if (calc->c->cos_engine) calc->c->cos_engine->tune.hkl_chunk = hkl_chunk;
    });
}

int dh_hc_set_options(dh_calculator* calc, unsigned flags)
{
    return guarded([&] {
        const unsigned all = DH_OPT_GENERAL_ROTATIONS | DH_OPT_RADIAL_RECORDS | DH_OPT_RADIAL_TABLES |
                             DH_OPT_DIRECT_PHASES;
        if (!calc || (flags & ~all) || ((flags & DH_OPT_RADIAL_RECORDS) && (flags & DH_OPT_RADIAL_TABLES)))
            throw Error(DH_E_ARGUMENT, "invalid options");
        for (Engine* e : {calc->c->engine.get(), calc->c->iam_engine.get(), calc->c->cos_engine.get()}) {
            if (!e) continue;
            e->tune.general_path = (flags & DH_OPT_GENERAL_ROTATIONS) != 0;
            e->tune.radial_mode = (flags & DH_OPT_RADIAL_RECORDS) ? 1 : (flags & DH_OPT_RADIAL_TABLES) ? 2 : 0;
            e->tune.direct_phases = (flags & DH_OPT_DIRECT_PHASES) != 0;
        }
    });
}

int dh_hc_upload(dh_calculator* calc, const dh_atoms* atoms, int n_hkl, const int* hkl, const double* dtarget_df)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        upload(calc->c, atoms, n_hkl, hkl, dtarget_df);
    });
}

int dh_hc_run(dh_calculator* calc, unsigned selector)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        calc->c->last_selector = selector & (DH_D_ALL | DH_D_MULTIPOLE);
        check_selector(calc->c, selector);
        Engine& e = calc->c->host_eng();
        e.run(calc->c->prec, calc->c->last_selector != 0, run_mp(calc->c, calc->c->last_selector));
        calc->c->last_run = &e;
    });
}

int dh_hc_fetch(dh_calculator* calc, double* f, double* d_xyz, double* d_adp, double* d_occ)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        fetch(calc->c, f, d_xyz, d_adp, d_occ);
    });
}

int dh_hc_sf_and_derivs(dh_calculator* calc, const dh_atoms* atoms, int n_hkl, const int* hkl,
                        const double* dtarget_df, unsigned selector, double* f, double* d_xyz, double* d_adp,
                        double* d_occ)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        upload(calc->c, atoms, n_hkl, hkl, dtarget_df);
        calc->c->last_selector = selector & (DH_D_ALL | DH_D_MULTIPOLE);
        check_selector(calc->c, selector);
        Engine& e = calc->c->host_eng();
        e.run(calc->c->prec, calc->c->last_selector != 0, run_mp(calc->c, calc->c->last_selector));
        calc->c->last_run = &e;
        fetch(calc->c, f, d_xyz, d_adp, d_occ);
    });
}

int dh_hc_set_reflections(dh_calculator* calc, int n_hkl, const int* hkl)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        set_reflections(calc->c, n_hkl, hkl);
    });
}

int dh_hc_sf_and_derivs_dev(dh_calculator* calc, const dh_atoms_dev* atoms, const double* dtarget_df,
                            unsigned selector, double* f, double* d_xyz, double* d_adp, double* d_occ, void* stream)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        sf_dev(calc->c, atoms, dtarget_df, selector, f, d_xyz, d_adp, d_occ, stream, false);
    });
}

int dh_hc_sf_and_derivs_mp_dev(dh_calculator* calc, const dh_atoms_dev* atoms, const double* dtarget_df,
                               unsigned selector, double* f, double* d_xyz, double* d_adp, double* d_occ,
                               double* d_plm, double* d_p_val, double* d_kappa, double* d_kappa_prime, void* stream)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        double* const tsum[4] = {d_plm, d_p_val, d_kappa, d_kappa_prime};
        sf_dev(calc->c, atoms, dtarget_df, selector, f, d_xyz, d_adp, d_occ, stream, false, true, tsum);
    });
}

int dh_hc_sf_and_derivs_dev_noalloc(dh_calculator* calc, const dh_atoms_dev* atoms, const double* dtarget_df,
                                    unsigned selector, double* f, double* d_xyz, double* d_adp, double* d_occ,
                                    void* stream)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        sf_dev(calc->c, atoms, dtarget_df, selector, f, d_xyz, d_adp, d_occ, stream, true);
    });
}

int dh_hc_autotune(dh_calculator* calc, unsigned selector, int* chunk)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        check_selector(calc->c, selector);
        int c = calc->c->host_eng().autotune(calc->c->prec, (selector & (DH_D_ALL | DH_D_MULTIPOLE)) != 0,
                                          run_mp(calc->c, selector));
        if (chunk) *chunk = c;
    });
}

int dh_hc_set_type_parameters(dh_calculator* calc, int n_types, const double* plm, const double* p_val,
                              const double* kappa, const double* kappa_prime)
{
    return guarded([&] {
        if (!calc || !plm || !p_val || !kappa || !kappa_prime) throw Error(DH_E_ARGUMENT, "null argument");
        Calculator* c = calc->c;
        if (n_types != (int)c->types.size()) throw Error(DH_E_ARGUMENT, "number of atom types differs from the model");
        std::vector<AtomType> t = c->types;
        for (int i = 0; i < n_types; i++) {
            for (int j = 0; j < 25; j++) t[i].plm[j] = plm[25 * i + j];
            t[i].p_val = p_val[i];
            t[i].kappa = kappa[i];
            t[i].kappa_prime = kappa_prime[i];
        }
        std::vector<AtomType> et(c->model.n_types);
        for (int e = 0; e < c->model.n_types; e++) et[e] = t[c->model.type_origin[e]];
        c->engine->set_types(et);
        if (c->cos_engine) c->cos_engine->set_types(et);
        c->types = t;
        c->model = c->engine->model();
    });
}

int dh_hc_fetch_type_derivatives(dh_calculator* calc, double* d_plm, double* d_p_val, double* d_kappa,
                                 double* d_kappa_prime)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        Calculator* c = calc->c;
        if (!c->have_mp) throw Error(DH_E_ARGUMENT, "no multipole derivatives: run with DH_D_MULTIPOLE and fetch first");
        const size_t nt = c->types.size();
        if (d_plm) std::memcpy(d_plm, c->d_plm.data(), 25 * nt * sizeof(double));
        if (d_p_val) std::memcpy(d_p_val, c->d_pval.data(), nt * sizeof(double));
        if (d_kappa) std::memcpy(d_kappa, c->d_kappa.data(), nt * sizeof(double));
        if (d_kappa_prime) std::memcpy(d_kappa_prime, c->d_kappa_prime.data(), nt * sizeof(double));
    });
}

int dh_hc_set_timing(dh_calculator* calc, int on)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        calc->c->engine->timing = on != 0;
        if (calc->c->iam_engine) calc->c->iam_engine->timing = on != 0;
        if (calc->c->cos_engine) calc->c->cos_engine->timing = on != 0;
    });
}

int dh_hc_last_timing(dh_calculator* calc, double* ms, int* gx, int* gy, int* chunk)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        Engine& e = calc->c->last_run ? *calc->c->last_run : calc->c->eng();
        if (ms) *ms = e.last_main_ms;
        if (gx) *gx = e.last_grid_x;
        if (gy) *gy = e.last_grid_y;
        if (chunk) *chunk = e.last_chunk;
    });
}

int dh_hc_engine_info(dh_calculator* calc, int* n_ops, int* order, int* centro, char* centering)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        if (n_ops) *n_ops = calc->c->eng().n_ops();
        if (order) *order = calc->c->sg.order;
        if (centro) *centro = calc->c->sg.centro ? 1 : 0;
        if (centering) *centering = calc->c->sg.centering;
    });
}

int dh_hc_per_hkl(dh_calculator* calc, const dh_atoms* atoms, int n_hkl, const int* hkl, dh_per_hkl_fn fn,
                  void* user)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        if (!fn) throw Error(DH_E_ARGUMENT, "null per-reflection callback");
        upload(calc->c, atoms, n_hkl, hkl, nullptr, true);
        per_hkl(calc->c, fn, user);
    });
}

int dh_hc_jacobian_upload(dh_calculator* calc, const dh_atoms* atoms, int n_hkl, const int* hkl)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        upload(calc->c, atoms, n_hkl, hkl, nullptr, true);
    });
}

int dh_hc_jacobian_rows(dh_calculator* calc, int k0, int k1, double* f, double* rows)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        rows_to_host(calc->c, k0, k1, f, rows);
    });
}

int dh_hc_jacobian_rows_dev(dh_calculator* calc, int k0, int k1, void* f_dev, void* rows_dev, void* stream)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        Calculator* c = calc->c;
        check_rows(c, k0, k1);
        if (k1 == k0) return;
        if (!f_dev || !rows_dev) throw Error(DH_E_ARGUMENT, "null output pointer");
        hipStream_t caller = static_cast<hipStream_t>(stream), st = c->eng().stream();
        if (!c->ev_in) HIP_CHECK(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming));
        if (!c->ev_out) HIP_CHECK(hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming));
        // the engine stream starts after the caller's earlier work and the caller's stream resumes
        // after the rows are written
        HIP_CHECK(hipEventRecord(c->ev_in, caller));
        HIP_CHECK(hipStreamWaitEvent(st, c->ev_in, 0));
        c->eng().jacobian(k0, k1, static_cast<double*>(f_dev), static_cast<double*>(rows_dev), c->prec);
        c->last_run = &c->eng();
        HIP_CHECK(hipEventRecord(c->ev_out, st));
        HIP_CHECK(hipStreamWaitEvent(caller, c->ev_out, 0));
    });
}

static void form_factors(Calculator* c, const double* lcs, const unsigned char* include, int n_h, const double* h_cart,
                         double* f)
{
    if (n_h < 0) throw Error(DH_E_ARGUMENT, "negative number of reflections");
    if (n_h > 0 && (!h_cart || !f)) throw Error(DH_E_ARGUMENT, "null argument");
    const int n = (int)c->atom_type.size();
    if (n_h == 0 || n == 0) return;
    std::vector<double> frames(9 * (size_t)n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 9; j++) frames[9 * (size_t)i + j] = lcs ? lcs[9 * (size_t)i + j] : (j % 4 == 0 ? 1.0 : 0.0);
    if (c->iam)   // HansenCoppensStructureFactorCalculator.cpp:1246-1247
        throw Error(DH_E_ARGUMENT, "using functionality available only for CPU calcluations of multipole model form "
                                   "factors");
    c->engine->form_factors(n, c->atom_type.data(), frames.data(), n_h, h_cart, f);
    c->last_run = c->engine.get();
    if (include)   // countAtom: excluded atoms get 0 (HansenCoppens_SF_Engine.cpp:467-471)
        for (int k = 0; k < n_h; k++)
            for (int i = 0; i < n; i++)
                if (!include[i]) f[2 * ((size_t)k * n + i)] = f[2 * ((size_t)k * n + i) + 1] = 0.0;
}

int dh_hc_form_factors(dh_calculator* calc, const double* lcs, const unsigned char* include, int n_h, const int* hkl,
                       double* f)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        if (n_h > 0 && !hkl) throw Error(DH_E_ARGUMENT, "null Miller indices");
        Calculator* c = calc->c;
        // h_cart = h a* + k b* + l c* (ReciprocalLatticeUnitCell::fractionalToCartesian)
        std::vector<double> hc(3 * (size_t)std::max(n_h, 0));
        for (int k = 0; k < n_h; k++)
            for (int j = 0; j < 3; j++)
                hc[3 * (size_t)k + j] = hkl[3 * k] * c->cell.c2f[j] + hkl[3 * k + 1] * c->cell.c2f[3 + j] +
                                        hkl[3 * k + 2] * c->cell.c2f[6 + j];
        form_factors(c, lcs, include, n_h, hc.data(), f);
    });
}

int dh_hc_form_factors_cart(dh_calculator* calc, const double* lcs, const unsigned char* include, int n_h,
                            const double* h_cart, double* f)
{
    return guarded([&] {
        if (!calc) throw Error(DH_E_ARGUMENT, "null calculator");
        form_factors(calc->c, lcs, include, n_h, h_cart, f);
    });
}

int dh_batch_create(dh_calculator* const* calcs, int n, dh_batch** out)
{
    return guarded([&] {
        if (!out || !calcs || n <= 0) throw Error(DH_E_ARGUMENT, "invalid batch arguments");
        *out = nullptr;
        std::unique_ptr<dh_batch> b(new dh_batch);
        std::vector<Engine*> eng;
        for (int i = 0; i < n; i++) {
            if (!calcs[i]) throw Error(DH_E_ARGUMENT, "null calculator in batch");
            if (calcs[i]->c->iam || calcs[i]->c->electron || calcs[i]->c->lcs_chain)
                throw Error(DH_E_ARGUMENT, "batches take multipole-model X-ray calculators (not CPU_IAM, electron "
                                           "or the local-frame chain rule)");
            b->calcs.push_back(calcs[i]->c);
            // the engine that holds the calculator's host upload (its coset engine where it has one)
            eng.push_back(&calcs[i]->c->host_eng());
        }
        b->engines = eng;
        b->engine.reset(new BatchEngine(eng));
        *out = b.release();
    });
}

void dh_batch_destroy(dh_batch* batch) { delete batch; }

static Precision batch_precision(dh_batch* b)
{
    const Precision p = b->calcs[0]->prec;
    for (size_t i = 0; i < b->calcs.size(); i++) {
        Calculator* c = b->calcs[i];
        if (c->prec != p) throw Error(DH_E_ARGUMENT, "calculators of a batch must share the calculation mode");
        if (c->iam || c->electron || c->lcs_chain)
            throw Error(DH_E_ARGUMENT, "batches take multipole-model X-ray calculators (not CPU_IAM, electron or the "
                                       "local-frame chain rule)");
        if (&c->host_eng() != b->engines[i])
            throw Error(DH_E_ARGUMENT, "a calculator of the batch changed its kernel options since the batch was "
                                       "created");
    }
    return p;
}

int dh_batch_run(dh_batch* batch, unsigned selector, int use_graph)
{
    return guarded([&] {
        if (!batch) throw Error(DH_E_ARGUMENT, "null batch");
        if (selector & ~(unsigned)DH_D_ALL) throw Error(DH_E_ARGUMENT, "multipole derivatives are not batched");
        const Precision p = batch_precision(batch);
        batch->engine->run(p, selector != 0, use_graph != 0);
        for (Calculator* c : batch->calcs) {
            c->last_selector = selector;
            c->have_mp = false;
        }
    });
}

int dh_batch_autotune(dh_batch* batch, unsigned selector, int* chunk)
{
    return guarded([&] {
        if (!batch) throw Error(DH_E_ARGUMENT, "null batch");
        if (selector & ~(unsigned)DH_D_ALL) throw Error(DH_E_ARGUMENT, "multipole derivatives are not batched");
        const int c = batch->engine->autotune(batch_precision(batch), selector != 0);
        if (chunk) *chunk = c;
    });
}

int dh_batch_set_tuning(dh_batch* batch, int hkl_chunk)
{
    return guarded([&] {
        if (!batch || hkl_chunk < 0) throw Error(DH_E_ARGUMENT, "invalid tuning");
        batch->engine->hkl_chunk = hkl_chunk;
    });
}

int dh_batch_last_timing(dh_batch* batch, double* step_ms, int* chunk, int* units, int* grid)
{
    return guarded([&] {
        if (!batch) throw Error(DH_E_ARGUMENT, "null batch");
        if (step_ms) *step_ms = batch->engine->last_ms;
        if (chunk) *chunk = batch->engine->last_chunk;
        if (units) *units = batch->engine->last_units;
        if (grid) *grid = batch->engine->last_grid;
    });
}

}  // extern "C"
