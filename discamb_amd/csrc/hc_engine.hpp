// Device engine of the Hansen-Coppens structure-factor path (one GPU, one HIP stream).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "hc_internal.hpp"

namespace discamb_amd {

#define HIP_CHECK(x)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess)                                                                     \
            throw Error(3, std::string("HIP error ") + hipGetErrorString(e_) + " at " #x);       \
    } while (0)

enum Precision { kFp64 = 0, kMixed = 1 };

struct BatchItem;      // per-structure arguments of the batched kernels (device layout)
struct MainArgs;

// One explicit operation as the engine uses it.
struct EngineOp {
    int rf[9];          // fractional rotation
    double tf[3];       // fractional translation
    double rc[9];       // Cartesian rotation
    double anom_sign;   // +1, or -1 for inversion partners (conjugated anomalous term)
};

struct DeviceBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n);
    void release();
    ~DeviceBuffer() { release(); }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Page-locked host staging for the per-call copies (grown on demand, kept by the engine).
struct HostBuffer {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n);
    void release();
    ~HostBuffer() { release(); }
    template <typename T> T* at(size_t byte_offset) const
    {
        return reinterpret_cast<T*>(static_cast<char*>(p) + byte_offset);
    }
};

// Device-resident calls: the caller's atom arrays, dT/dF and outputs in device memory, in the
// caller's order and conventions (the conversions run on the device).
struct DevCall {
    const double* xyz = nullptr;             // 3n
    const int* adp_size = nullptr;           // n: 1 (U_iso) or 6; any other value: no ADP
    const double* adp = nullptr;             // 6n
    const double* occ = nullptr;             // n
    const double* lcs = nullptr;             // 9n, or null: identity frames
    const double* anom = nullptr;            // 2n, or null: the wavefunction types' f', f''
    const unsigned char* include = nullptr;  // n, or null: every atom
    const double* dtdf = nullptr;            // 2 n_hkl: dT/dA + i dT/dB (not conjugated)
    double* f = nullptr;                     // 2 n_hkl
    double* d_xyz = nullptr;                 // 3n, 6n, n (each may be null)
    double* d_adp = nullptr;
    double* d_occ = nullptr;
    unsigned selector = 0;                   // xyz 1, ADP 2, occupancy 4
    // multipole-parameter derivatives (the two adjoint passes), summed per calculator type:
    // 25 n_types, n_types, n_types, n_types (each may be null)
    bool multipole = false;
    double* d_plm = nullptr;
    double* d_pval = nullptr;
    double* d_kappa = nullptr;
    double* d_kappa_prime = nullptr;
};

// Conversions between the caller's conventions and the engine's (fixed per calculator).
struct DevConv {
    int xyz_identity = 1;    // fractional input coordinates (copied)
    double c2f[9];           // otherwise Cartesian -> fractional
    double u2c[36];          // ADP input -> 2 pi^2 U_cart (row-major)
    int dxyz_identity = 1;   // fractional output derivatives
    double dxyz_m[9];        // otherwise caller xyz derivative = dxyz_m . d/dx_frac
    double dadp_m[36];       // caller ADP derivative = dadp_m . d/dU_cart
    int centric_origin = 0;  // dT/dF -> (2 Re, 0) (F = 2 Re of the half sum)
};

struct TuneParams {
    int hkl_chunk = 0;       // reflections per workgroup column; 0 = automatic
    bool general_path = false;  // force the general-rotation kernels for diagonal groups (tests)
    int radial_mode = 0;        // 0 automatic, 1 per-(reflection, type) records, 2 on-the-fly tables
    bool direct_phases = false; // FP64: evaluate every phase and Debye-Waller factor directly (tests)
};

// Atom data for one call, in the order the caller supplies them (already converted to the
// engine conventions: fractional xyz, U_cart * 2 pi^2 as a 6-vector, iso atoms as U I).
struct EngineAtoms {
    int n = 0;
    std::vector<double> xyz_frac;   // 3n
    std::vector<double> u6;         // 6n
    std::vector<int> iso;           // n
    std::vector<double> weight;     // occ * mult / |G|
    std::vector<double> anom;       // 2n (per-atom f', f'')
    std::vector<double> lcs;        // 9n, row-major, columns = local axes
    std::vector<int> type;          // n
    std::vector<unsigned char> include;
};

// Per-reflection derivative rows (the per-hkl path): the caller's derivative conventions and the
// per-atom data the rows need, in caller atom order.
struct JacConv {
    double xyz_m[9];              // caller xyz derivative = xyz_m . d/dx_frac (row-major)
    double adp_m[36];             // caller ADP derivative = adp_m . d/dU_cart (row-major)
    std::vector<double> mult_w;   // m / |G|
    std::vector<int> adp_kind;    // 0 (no ADP), 1 (U_iso), 6
};

class Engine {
public:
    Engine(int device, const EngineModel& model, const double* cell6, const std::vector<EngineOp>& ops,
           bool centric_origin);
    ~Engine();
    // Upload one call's inputs; hkl as int triples, dtdf already conj * centering multiplier
    // followed by the n_hkl centering multipliers.  By default reflections are processed in order
    // of increasing |h| and atoms grouped by multipole class and type (results are returned in
    // caller order either way); rows_layout keeps the caller's order of both on the device, which
    // the per-reflection rows need.
    void upload(const EngineAtoms& atoms, int n_hkl, const int* hkl, const double* dtdf, bool rows_layout = false);
    // Forget the inputs of the last upload (and its results): run() then refuses until the next
    // upload completes.  The reflection-dependent prep stays valid for its reflection list.
    void invalidate_inputs()
    {
        uploaded_ = false;
        ran_ = false;
        jac_ok_ = false;
        dev_ready_ = false;
    }
    // Per-reflection derivative rows after an upload with rows_layout: set the conventions,
    // then compute F and the rows of reflections [k0, k1) into device memory, asynchronously on
    // the engine stream.  rows: [k - k0][10 planes][n_atoms] complex; f: [k - k0] complex.
    void set_jacobian_conventions(const JacConv& conv);
    // prec: kFp64, or kMixed for the mixed-precision terms (FP64 rows either way)
    void jacobian(int k0, int k1, double* f_dev, double* rows_dev, Precision prec = kFp64);
    int n_atoms() const { return n_atoms_; }
    // Electron scattering (Mott-Bethe) for the summed path: z_f0 = (Z, f_e(0)) per atom type.
    // The radial factors then come from per-(reflection, type) records.
    void set_electron(bool on, const std::vector<double>& z_f0);
    bool electron() const { return electron_; }
    // With the multipole adjoint: also dT/dL (9 per atom, local axes as columns, atom weight not
    // applied) for the chain rule through recomputed local frames; read with download_lcs_grad.
    bool lcs_grad = false;
    // Order reflections along hkl lines at the next upload (set for the FP64 drop-in, whose line
    // kernels carry phase and Debye-Waller factors along them); otherwise by increasing |h|.
    bool line_order = false;
    // With the multipole adjoint: whether the Pval / kappa / kappa' sums (and the P_lm derivatives)
    // are wanted as well; without them only the B_lm pass is added to the plain one.
    bool mp_scalars = true;
    void download_lcs_grad(double* g);
    // Form-factor route: f[k][atom] (complex, caller atom order) for n_h Cartesian h vectors and
    // the given local frames and atom types; synchronous, host pointers.
    void form_factors(int n_atoms, const int* type, const double* lcs, int n_h, const double* h_cart, double* f);
    // Run all kernels on the uploaded inputs (asynchronous on the engine stream).  mp adds the
    // multipole-parameter adjoint (implies derivs).
    void run(Precision prec, bool derivs, bool mp = false);
    // Copy F (2 n_hkl) and per-atom raw derivatives back, in caller atom order; synchronises.
    // Per atom nd() values: dxyz_frac[3] (x 2 pi w), dU_cart[6] (x 2 pi^2 w), docc (x mult/|G|),
    // and with the multipole adjoint dP_lm local frame [25], dPval, dkappa, dkappa' (all x w).
    void download(double* f, double* d);
    int nd() const { return last_nd_; }
    // Replace the atom-type parameters (P_lm, Pval, kappa, kappa'); the multipole truncation
    // max L is recomputed as the reference does on every call.
    void set_types(const std::vector<AtomType>& types);
    const EngineModel& model() const { return model_; }
    void synchronize();
    // Time the fused kernel plus its stage-2 reductions over candidate reflection-chunk sizes on
    // the uploaded inputs and keep the fastest for this (atoms, reflections, ops, precision)
    // shape; returns the chunk chosen.
    int autotune(Precision prec, bool derivs, bool mp = false);
    hipStream_t stream() const { return stream_; }
    TuneParams tune;
    // last run: launch geometry of the main kernel
    int last_grid_x = 0, last_grid_y = 0, last_chunk = 0;
    float last_main_ms = 0.f;     // main-kernel time measured with events when timing_ is on
    bool timing = false;
    int n_ops() const { return (int)ops_.size(); }
    int device() const { return device_; }
    int n_hkl() const { return n_hkl_; }
    int n_slots() const { return n_sorted_; }
    // Batched steps (BatchEngine): make every buffer of a run exist and describe this structure's
    // part of the step; after the batched launch, mark the results as a completed run.
    void batch_item(Precision prec, bool derivs, int chunk, BatchItem* out);
    void mark_batched_run(bool derivs);
    // changes whenever the device reflection list (its content or order) changes
    unsigned long long hkl_generation() const { return hkl_gen_; }
    int tab_nodes(int wfn) const { return wfn < (int)tab_nodes_.size() ? tab_nodes_[wfn] : 3; }
    // Device-resident calls.  prepare_dev uploads the reflection list (host memory) with every
    // atom of the model in its slot (excluded atoms get weight 0 per call), keeps the reflection
    // order, centring multipliers (mult, caller order) and per-atom constants (m/|G|, default
    // f' + i f'') on the device, and prepares the reflection records for prec.  run_dev then runs
    // one call on device pointers, ordered on `stream`, without host synchronisation.  With
    // scratch != null every per-call buffer (atom records, sorted dT/dF, partial sums, sorted F,
    // raw derivatives) is carved out of that memory and nothing is allocated: a call that would
    // need more memory than scratch_bytes, or reflection records of another calculation mode,
    // is refused.  Calls sharing one scratch must not overlap.
    void prepare_dev(const EngineAtoms& layout, int n_hkl, const int* hkl, const double* mult,
                     const std::vector<double>& mult_w, const std::vector<double>& anom_default,
                     const DevConv& conv, Precision prec);
    void run_dev(const DevCall& call, Precision prec, hipStream_t stream, void* scratch, size_t scratch_bytes);
    // the calculator type of every atom and each type's multipole order + 1 (P_lm entries with
    // l >= n_l are not parameters): the per-type sums of the device calls' multipole derivatives
    void set_type_sum_map(const std::vector<int>& atom_type, const std::vector<int>& n_l);
    // device-resident calls over the images of the caller's atoms under coset representatives:
    // prepare_dev laid out K images per atom (image k of atom i at i K + k); run_dev builds them
    // from the caller's arrays and folds their derivatives back.  Empty: no images.
    void set_dev_cosets(const std::vector<EngineOp>& reps);
    bool dev_ready() const { return dev_ready_; }
    // the static scratch array of the allocation-free calls on this engine's device, its size
    void* scratch() const { return scratch_; }
    static size_t scratch_bytes();

private:
    void upload_checked(const EngineAtoms& atoms, int n_hkl, const int* hkl, const double* dtdf, bool rows_layout);
    void ensure_prep_buffers(Precision prec, bool mp);
    void plan_tables();
    void main_args(Precision prec, bool derivs, bool mp, int chunk, MainArgs* a);
    // general: the per-(reflection, operation) records even where the parity-folded ones apply
    void run_prep(Precision prec, bool mp, bool electron = true, bool general = false);
    void launch_main(Precision prec, bool derivs, bool mp, int chunk);
    void reduce(bool derivs, bool mp, int chunk);
    void fill(DeviceBuffer& b, size_t n, double v);
    void build_splines();
    void fill(double* p, size_t n, double v);
    int jac_chunk(int n_rows, Precision prec);
    int resident_waves(bool mixed, bool derivs, bool mp);
    int default_chunk(bool mixed, bool derivs, bool mp);
    int chunk_for(Precision prec, bool derivs, bool mp);
    long long shape_key(Precision prec, bool derivs, bool mp) const;
    int type_class(int t) const;
    // parity-folded kernels: mixed precision only (the FP64 variant needs twice the registers for
    // its class sums and loses more to occupancy than it saves)
    bool use_diag(bool mixed, bool mp) const { return diag_ && mixed && !tune.general_path; }
    // radial factors evaluated per (atom, reflection) from the per-wavefunction tables instead of
    // per-(reflection, type) records: chosen when types are not shared by many atoms (the records
    // would then cost more to write and read than the lookups), never with the multipole adjoint
    bool use_otf(bool mp) const
    {
        if (mp || electron_) return false;
        if (tune.radial_mode != 0) return tune.radial_mode == 2;
        return (long long)model_.n_types * 8 > (long long)n_sorted_;
    }
    int resident_waves_[32 * 7] = {};   // key: kernel kind (0-6) * 32 + path bits
    bool diag_ = false;               // every operation is a diagonal sign matrix
    int diag_pat_[8] = {};
    float diag_anom_[8] = {};
    int last_nd_ = 10;
    std::map<long long, int> tuned_;
    int device_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    hipEvent_t ev_up_ = nullptr;    // recorded after the copies of the last upload
    bool up_pending_ = false;       // ... which may still read the staging buffer
    // inputs of the atom order of the last upload (unchanged: the order is kept)
    bool last_rows_layout_ = false;
    std::vector<int> last_atom_type_, wave_cls_;
    std::vector<unsigned char> last_include_;
    EngineModel model_;
    std::vector<EngineOp> ops_;
    bool centric_origin_;
    double f2c_[9], c2f_[9];
    int n_atoms_ = 0, n_sorted_ = 0, n_hkl_ = 0;
    bool ran_ = false;
    bool uploaded_ = false;         // the last upload completed (run() needs one)
    bool part_f32_ = false;         // the F partials of the last main launch are FP32 (mixed precision)
    bool dev_ready_ = false;        // prepare_dev set the layout of device-resident calls
    int chunk_override_ = 0;        // reflection chunk forced by an allocation-free call
    DevConv dev_conv_;
    DeviceBuffer d_dev_perm_, d_dev_mult_, d_dev_atomc_;
    DeviceBuffer d_tsum_map_;       // set_type_sum_map: atoms by type, type offsets, n_l per type
    DeviceBuffer d_dev_reps_;       // set_dev_cosets: 21 doubles per representative
    int dev_K_ = 1;
    int tsum_types_ = 0;
    void* scratch_ = nullptr;       // device address of the static scratch array
    std::vector<int> order_;        // sorted slot -> caller atom index (-1 padding)
    std::vector<int> hkl_perm_;     // device reflection position -> caller reflection index
    std::vector<int> last_hkl_;     // reflections of the last upload (unchanged lists skip the sort)
    int last_order_ = -1;           // reflection order of the last upload: 0 caller, 1 by |h|, 2 lines
    // FP64 line recurrences: the axis of the lines, per reflection whether it continues the line
    // of the one before, per operation [Delta R_f (3), Delta . t, m(Delta_s) (6)]
    int line_axis_ = 2;
    DeviceBuffer d_line_cont_, d_line_ops_;
    std::vector<int> line_cont_host_;
    std::vector<double> line_ops_host_;
    bool line_ops_ok() const { return !ops_.empty() && ops_.size() <= 4; }
    void line_sort(int n_hkl, const int* hkl, std::vector<std::pair<double, int>>& key);
    bool use_line(bool mixed, bool mp) const
    {
        if (tune.direct_phases || last_order_ != 2 || !line_ops_ok() || use_otf(mp)) return false;
        // the mixed parity-folded line kernels exist for the compiled sign-pattern sets only, and
        // carry the multipole passes too; the other line kernels run without them
        if (mp) return diag_ && !tune.general_path && diag64_code() != 0;
        return !use_diag(mixed, mp) || diag64_code() != 0;
    }
    // the FP64 line kernel with per-reflection records for diagonal groups: its sign-pattern code, 0
    // when the operations' patterns are not one of the compiled sets
    int diag64_code() const;
    // sign-pattern code passed to main_kernel: use_diag64, or for the mixed parity-folded kernel
    // with per-atom table lookups (no line recurrences) a compiled set marked kDiagDirect
    int main_code(bool mixed, bool mp) const;
    int use_diag64(bool mixed, bool mp) const
    {
        return use_line(mixed, mp) && diag_ && !tune.general_path ? diag64_code() : 0;
    }
    bool prep_valid_ = false;       // the prep buffers hold the records of prep_key_ (run_prep)
    unsigned long long hkl_gen_ = 0;
    int prep_key_ = -1;
    std::vector<int> tab_nodes_;    // radial-table node count per wavefunction
    // model tables
    DeviceBuffer d_core_, d_val_, d_type_, d_ops_;
    // per call
    DeviceBuffer d_hkl_, d_dtdf_, d_atoms_in_, d_atoms_, d_hklop_, d_hklop32_, d_phase_, d_radial_,
        d_radial32_, d_ht_, d_fpart_, d_dpart_, d_f_, d_d_, d_order_, d_radx_, d_tab_, d_tab32_, d_tabrec_;
    HostBuffer h_up_, h_down_;      // pinned staging of upload() and download()
    // per-reflection rows
    bool jac_ok_ = false;           // conventions set for the current upload
    bool jac_prepped_ = false;      // the prep tables hold the general records of the current upload
    Precision jac_prec_ = kFp64;    // ... in this precision
    bool rows_layout_ = false;
    int n_included_ = 0;
    int jac_resident_[2] = {0, 0}; // resident waves of the FP64 / mixed rows kernel
    DeviceBuffer d_jatoms_, d_jconv_, d_jrec_;
    bool electron_ = false;
    DeviceBuffer d_ez_, d_lcsg_;
    // form-factor route
    bool splines_ready_ = false;
    DeviceBuffer d_spline_, d_ff_ain_, d_ff_atoms_, d_ff_h_, d_ff_rec_, d_ff_hrec_, d_ff_out_;
    double h_max_ = 0.0;            // largest |h| of the uploaded reflections
    int n_tab_nodes_ = 0;           // form-factor table nodes over all wavefunctions
};

// Many structures in one step: the per-structure prep kernels, the fused F + adjoint kernel and
// the stage-2 reductions each run as ONE launch over all structures (flattened index spaces; the
// main kernel is persistent and takes (structure, atom block, reflection chunk) units from a
// device work counter).  Each structure keeps its own Engine (inputs uploaded with
// Engine::upload, results read with Engine::download); the batch only borrows them.  All
// engines must live on one device.
class BatchEngine {
public:
    explicit BatchEngine(const std::vector<Engine*>& engines);
    ~BatchEngine();
    // One batched step on the engines' uploaded inputs; synchronous.  graph: the step is captured
    // into a hipGraph on first use and replayed while the structures' buffers and launch shape
    // stay the same.
    void run(Precision prec, bool derivs, bool graph);
    // Times the whole batched step over candidate reflection chunks; keeps the fastest.
    int autotune(Precision prec, bool derivs);
    int hkl_chunk = 0;             // fixed reflection chunk for every structure; 0 = autotuned/default
    float last_ms = 0.f;           // device time of the last step (events around it)
    int last_chunk = 0, last_units = 0, last_grid = 0;

private:
    void plan(Precision prec, bool derivs, int chunk);
    void share_tables(Precision prec, const std::vector<int>& order, std::vector<BatchItem>& items);
    void launch(Precision prec, bool derivs);
    void launch_hkl_prep();
    int chunk_for(Precision prec, bool derivs);
    void drop_graph();
    // per-reflection records of every structure, owned by the batch: rebuilt only when a
    // structure's reflection list, the precision or the buffer changes (rec_key_)
    DeviceBuffer d_rec_;
    std::vector<unsigned long long> rec_key_, last_rec_key_;
    std::vector<Engine*> eng_;
    int device_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    // the kernel bodies' persistent launches run concurrently: body group g > 0 on side_[g - 1],
    // forked from stream_ after the prep (ev_fork_) and joined back before the reductions (ev_join_)
    hipStream_t side_[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork_ = nullptr, ev_join_[3] = {nullptr, nullptr, nullptr};
    DeviceBuffer d_items_, d_off_, d_counter_;
    // radial tables shared by the structures of the batch (one per distinct wavefunction)
    DeviceBuffer d_gwfn_, d_gcore_, d_gval_, d_gtabs_, d_ltabs_, d_gtab_, d_gtab32_;
    int n_shared_nodes_ = 0, n_shared_wfn_ = 0;
    std::vector<unsigned char> items_;     // host image of the BatchItem array
    std::vector<int> off_;                 // prefix offsets of every index space
    std::vector<int> total_;
    std::vector<unsigned char> graph_key_; // items_ + off_ + precision at capture time
    hipGraphExec_t exec_ = nullptr;
    std::map<long long, int> tuned_;
    int grid_[4][4] = {};                  // persistent main-kernel grid per (precision, derivatives, body)
    // structures [b0, b1) of the batch order run kernel body `body` as one persistent launch
    struct BodyGroup {
        int body, b0, b1, units;
    };
    std::vector<BodyGroup> groups_;
};

}  // namespace discamb_amd
