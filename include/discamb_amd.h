/*
 * discamb_amd - C ABI of the MI355X (gfx950) Hansen-Coppens structure-factor engine.
 *
 * Drop-in boundary for DiSCaMB's
 *   HansenCoppensStructureFactorCalculator(const Crystal&, const HC_ModelParameters&)
 *     include/discamb/Scattering/HansenCoppensStructureFactorCalculator.h:42
 *   ::setCalculationMode(const std::string&)                    src/.../HansenCoppensStructureFactorCalculator.cpp:350-372
 *   ::calculateStructureFactorsAndDerivatives(atoms, lcs, hkl, f, dTarget_dparam, dTarget_df,
 *       countAtomContribution[, DerivativesSelector])            .h:148-165, .cpp:380-535
 * and of the engine FFI it dispatches to for engine "GPU",
 *   hc_sf_gpu::calculate_multipolar_structure_factors_gpu        src/GPU/cuda/HC_SF_GPU.h:10-69
 * All arrays are plain host pointers in row-major order; nothing C++ crosses the boundary.
 * Every function returns 0 on success or a positive DH_E_* code; dh_last_error() then returns a
 * thread-local message.  One calculator object must not be used by two threads at once (as the
 * reference calculator, whose calculate methods mutate internal state).
 */
#ifndef DISCAMB_AMD_H
#define DISCAMB_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define DH_OK 0
#define DH_E_ARGUMENT 1   /* inconsistent sizes / indices (reference: on_error::throwException) */
#define DH_E_MODEL 2      /* invalid model or space group */
#define DH_E_DEVICE 3     /* HIP runtime failure */
#define DH_E_MEMORY 4     /* device allocation failure */
#define DH_E_NO_DEVICE 5  /* no usable GPU */
#define DH_E_CALLBACK 6   /* a per-reflection callback asked to stop (non-zero return) */

/* engine precision (setCalculationMode strings "HIP" and "HIP_MIXED") */
#define DH_PRECISION_FP64 0
#define DH_PRECISION_MIXED 1

/* DerivativesSelector bits (include/discamb/Scattering/SF_CalcDataTypes.h:44-56) */
#define DH_D_XYZ 1u
#define DH_D_ADP 2u
#define DH_D_OCC 4u
#define DH_D_ALL 7u
/* multipole-model parameters of the atom types: P_lm (local frame), Pval, kappa, kappa'.  No
 * counterpart in the reference, whose derivatives stop at xyz / ADP / occupancy. */
#define DH_D_MULTIPOLE 8u

/* xyz / ADP conventions (structural_parameters_convention) */
#define DH_XYZ_FRACTIONAL 0
#define DH_XYZ_CARTESIAN 1
#define DH_ADP_U_CIF 0
#define DH_ADP_U_CART 1

typedef struct dh_calculator dh_calculator;

/* Crystal + HC_ModelParameters, flattened.  Mirrors Crystal (unit cell, space group, atom
 * multiplicities, conventions) and HC_ModelParameters (wfn_parameters, type_parameters,
 * atom_to_wfn_map, atom_to_type_map; include/discamb/HC_Model/HC_ModelParameters.h:22-28). */
typedef struct {
    double cell[6];                 /* a, b, c (Angstrom), alpha, beta, gamma (degrees) */
    int n_symops;                   /* the full space group, as listed in a CIF */
    const int* symop_rot;           /* 9 * n_symops, fractional rotation, row-major */
    const double* symop_trans;      /* 3 * n_symops, fractional translation */
    int xyz_convention;             /* input coordinates: DH_XYZ_* */
    int adp_convention;             /* input ADPs: DH_ADP_* */
    int d_xyz_convention;           /* output derivatives */
    int d_adp_convention;
    /* wavefunction types (HC_WfnType, HC_WfnData.h:126-163) */
    int n_wfn;
    const int* wfn_orbital_offset;  /* n_wfn + 1, into the orbital arrays */
    const int* orbital_term_offset; /* n_orbitals + 1, into the Slater-term arrays */
    const int* orbital_nl;          /* 2 * n_orbitals: principal, azimuthal number */
    const double* term_coeff;       /* normalised Slater coefficients */
    const double* term_exp;         /* exponents, 1/Angstrom */
    const int* term_pow;            /* powers of r */
    const int* orbital_occupancy;   /* per orbital */
    const int* orbital_kind;        /* per orbital: 0 core, 1 valence, -1 unused */
    const double* def_val_exponent; /* n_wfn, zeta (1/Angstrom) */
    const int* def_val_pow_offset;  /* n_wfn + 1, into def_val_pow */
    const int* def_val_pow;         /* n_l */
    const double* anomalous;        /* 2 * n_wfn: f', f'' */
    /* atom types (HC_AtomTypeParameters, HC_AtomTypeParameters.h:19-50) */
    int n_types;
    const int* plm_n_l;             /* n_types: size of p_lm (number of l values, <= 5) */
    const double* plm;              /* 25 * n_types, index l*l + l + m, density normalised */
    const double* p_val;
    const double* kappa;            /* kappa_spherical_valence */
    const double* kappa_prime;      /* kappa_deformation_valence */
    /* atoms */
    int n_atoms;
    const int* atom_to_wfn;
    const int* atom_to_type;
    const double* atom_multiplicity;  /* AtomInCrystal::multiplicity (|G| for a general position) */
    /* optional (may be NULL): HC_WfnType::wfnAtomTypeName and ::atomic_number per wavefunction
     * type; the labels select the Gaussian form factors of the CPU_IAM mode, the atomic numbers
     * are needed for electron scattering */
    const char* const* wfn_label;
    const int* wfn_atomic_number;
} dh_model_desc;

/* Per-call atom parameters (vector<AtomInCrystal> + vector<Matrix3d> local frames). */
typedef struct {
    int n_atoms;
    const double* xyz;              /* 3n, in the model's xyz_convention */
    const int* adp_size;            /* n: 0, 1 (U_iso) or 6 (U11 U22 U33 U12 U13 U23) */
    const double* adp;              /* 6n, first adp_size[i] entries used */
    const double* occupancy;        /* n */
    const double* lcs;              /* 9n, row-major 3x3 with local x, y, z axes as columns */
    const double* anomalous;        /* 2n per-atom f', f'' or NULL (use the wavefunction type's) */
    const unsigned char* include;   /* n, countAtomContribution, or NULL (all) */
} dh_atoms;

/* Outputs: f is 2*n_hkl (re, im); derivatives follow TargetFunctionAtomicParamDerivatives
 * (SF_CalcDataTypes.h:36-41): d_xyz 3n, d_adp 6n (first adp_size[i] used), d_occ n.  Any
 * derivative pointer may be NULL. */

const char* dh_last_error(void);
const char* dh_version(void);
int dh_device_count(int* n);

/* HansenCoppensStructureFactorCalculator(crystal, parameters) on GPU `device`. */
int dh_hc_create(const dh_model_desc* model, int device, dh_calculator** out);
void dh_hc_destroy(dh_calculator* calc);
/* setCalculationMode (HansenCoppensStructureFactorCalculator.cpp:350-372): "HIP" (FP64 drop-in;
 * "GPU", "CPU" and "CPU_v4" are accepted as aliases so reference configuration strings keep
 * working - the computation always runs on the GPU), "HIP_MIXED", and the independent-atom model
 * "CPU_IAM" (alias "HIP_IAM"; "HIP_IAM_MIXED" for mixed precision): Waasmaier-Kirfel Gaussian form
 * factors of the wavefunction labels plus f' + i f'', no multipoles, with f'' unconjugated for
 * inversion partners as the reference's calculateSF_IAM (HansenCoppens_SF_Engine.cpp:229-278). */
int dh_hc_set_calculation_mode(dh_calculator* calc, const char* mode);
/* setElectronScattering (HansenCoppensStructureFactorCalculator.h:63, .cpp:510-514): multipole-model
 * structure factors for electrons by the Mott-Bethe formula, f_e = 0.023934 (Z - f_x) / (|h|^2 / 4)
 * + f' + i f'' and the "electron-IT" Gaussian value at h = 0 (HansenCoppens_SF_Engine3.cpp:1387-1398).
 * Needs dh_model_desc.wfn_atomic_number.  As in the reference it affects the summed path only (not
 * the per-reflection rows, the form factors or CPU_IAM); batches do not take it. */
int dh_hc_set_electron_scattering(dh_calculator* calc, int on);
/* N-Gaussian table of the independent-atom model: "Waasmeier-Kirfel" (default, the table of
 * CPU_IAM), "IT92" (AnyIamCalculator's default, AnyIamCalculator.cpp:101-108), "electron-IT" or
 * "electron-cctbx" (n_gaussian_form_factors_table, NGaussianFormFactorsTable.cpp).  In IAM mode
 * the terms are summed over every symmetry operation with f'' unconjugated, which is also what
 * AnyScattererStructureFactorCalculator computes (AnyScattererStructureFactorCalculator.cpp:
 * 614-832), so an IAM calculator with the "IT92" table stands in for AnyIamCalculator. */
int dh_hc_set_iam_table(dh_calculator* calc, const char* table);
/* Launch tuning: reflections per workgroup column (0 = automatic). */
int dh_hc_set_tuning(dh_calculator* calc, int hkl_chunk);
/* Kernel-path options (no reference counterpart; for tests and diagnosis).
 * DH_OPT_GENERAL_ROTATIONS: use the general-rotation kernels even when every symmetry operation is
 * a diagonal sign matrix (the engine otherwise folds such operations through harmonic parities). */
#define DH_OPT_GENERAL_ROTATIONS 1u
/* DH_OPT_RADIAL_RECORDS / DH_OPT_RADIAL_TABLES: radial factors from per-(reflection, atom type)
 * records, or evaluated per (atom, reflection) from per-wavefunction tables; by default the
 * engine takes tables when atom types are not shared by at least eight atoms on average. */
#define DH_OPT_RADIAL_RECORDS 2u
#define DH_OPT_RADIAL_TABLES 4u
/* DH_OPT_DIRECT_PHASES: every term evaluates its phase factor and Debye-Waller factor directly,
 * with reflections ordered by |h| (by default groups with at most four explicit operations order
 * reflections along hkl lines and carry both factors along them, as the reference's Engine3 does,
 * scattering_utilities.cpp:520-756); takes effect at the next upload. */
#define DH_OPT_DIRECT_PHASES 8u
/* Groups with a diagonal subgroup beyond it (tetragonal, hexagonal) run their summed calls on a
 * second engine over coset images unless DH_OPT_GENERAL_ROTATIONS is set; an upload or
 * dh_hc_set_reflections goes to the engine the options select at that moment, so after a change of
 * DH_OPT_GENERAL_ROTATIONS (or of the IAM table) dh_hc_run and the device-resident calls return an
 * error until the inputs or the reflections are set again. */
int dh_hc_set_options(dh_calculator* calc, unsigned flags);

/* calculateStructureFactorsAndDerivatives: synchronous, host pointers. hkl is 3*n_hkl int,
 * dtarget_df 2*n_hkl (dT/dA + i dT/dB, not conjugated). selector: DH_D_* bits; 0 = F only. */
int dh_hc_sf_and_derivs(dh_calculator* calc, const dh_atoms* atoms, int n_hkl, const int* hkl,
                        const double* dtarget_df, unsigned selector, double* f, double* d_xyz,
                        double* d_adp, double* d_occ);

/* The same call split in three for callers that keep inputs resident on the GPU between steps
 * (refinement loops, benchmarks): upload inputs, run the kernels, fetch results. */
int dh_hc_upload(dh_calculator* calc, const dh_atoms* atoms, int n_hkl, const int* hkl,
                 const double* dtarget_df);
int dh_hc_run(dh_calculator* calc, unsigned selector);
int dh_hc_fetch(dh_calculator* calc, double* f, double* d_xyz, double* d_adp, double* d_occ);
/* Replace the atom-type parameters (HC_AtomTypeParameters: p_lm 25 per type in the l*l + l + m
 * layout, p_val, kappa, kappa') between calls; the model's other data are kept. */
int dh_hc_set_type_parameters(dh_calculator* calc, int n_types, const double* plm, const double* p_val,
                              const double* kappa, const double* kappa_prime);
/* After a run with DH_D_MULTIPOLE and dh_hc_fetch: dT/d(type parameter), summed over the atoms of
 * each type; d_plm is 25 per type (entries beyond the type's p_lm size are 0). */
int dh_hc_fetch_type_derivatives(dh_calculator* calc, double* d_plm, double* d_p_val, double* d_kappa,
                                 double* d_kappa_prime);
/* Autotuner: on the uploaded inputs, time the fused kernel for candidate launch shapes (reflection
 * chunk per workgroup column, derived from the device's resident-wave capacity) in the current
 * calculation mode and keep the fastest for this workload shape.  Returns the chosen chunk. */
int dh_hc_autotune(dh_calculator* calc, unsigned selector, int* chunk);
/* Main-kernel time of the last dh_hc_run in ms (HIP events on the engine stream) and launch
 * geometry; enable with dh_hc_set_timing(calc, 1). */
int dh_hc_set_timing(dh_calculator* calc, int on);
int dh_hc_last_timing(dh_calculator* calc, double* main_kernel_ms, int* grid_x, int* grid_y, int* chunk);
/* explicit operations used by the engine (after inversion / centering folding) */
int dh_hc_engine_info(dh_calculator* calc, int* n_explicit_ops, int* group_order, int* centrosymmetric,
                      char* centering);

/* ---- Device-resident calls: calculateStructureFactorsAndDerivatives (.h:148-165, .cpp:380-508) on
 * device memory, for callers that keep their parameters on the GPU (torch tensors, refinement
 * loops, hipGraph capture).  The reference has no device-memory form; its GPU engine takes host
 * vectors (hc_sf_gpu::calculate_multipolar_structure_factors_gpu, src/GPU/cuda/HC_SF_GPU.h:10-69).
 * The reflection list is set once from host memory; it fixes the reflection order, the centring
 * multipliers and the reflection records, which every later device call on this calculator
 * reuses.  A host call or dh_hc_upload on the calculator replaces it (set it again); after a change
 * of calculation mode the device call prepares what the new mode needs on its first call. */
int dh_hc_set_reflections(dh_calculator* calc, int n_hkl, const int* hkl);
/* Atom parameters in device memory, layouts as dh_atoms (n_atoms = the model's atom count).
 * adp_size: 1 (U_iso) or 6; any other value means no ADP.  lcs, anomalous, include may be NULL
 * (identity frames, the wavefunction types' f' + i f'', every atom). */
typedef struct {
    int n_atoms;
    const double* xyz;
    const int* adp_size;
    const double* adp;
    const double* occupancy;
    const double* lcs;
    const double* anomalous;
    const unsigned char* include;
} dh_atoms_dev;
/* One call on device pointers, ordered on `stream` (a hipStream_t; NULL = the null stream) and
 * returning without waiting on the host.  dtarget_df: 2 n_hkl device doubles (dT/dA + i dT/dB, not
 * conjugated; may be NULL when selector is 0); f: 2 n_hkl; d_xyz 3n, d_adp 6n, d_occ n device
 * doubles (any may be NULL), in the calculator's derivative conventions.  selector: DH_D_XYZ |
 * DH_D_ADP | DH_D_OCC bits (DH_D_MULTIPOLE: dh_hc_sf_and_derivs_mp_dev; local frames recomputed
 * from positions need the host call).  Results equal the host call's to rounding (the convention
 * changes run on the device).
 * Argument errors are reported synchronously; device-side values are not validated. */
int dh_hc_sf_and_derivs_dev(dh_calculator* calc, const dh_atoms_dev* atoms, const double* dtarget_df,
                            unsigned selector, double* f, double* d_xyz, double* d_adp, double* d_occ, void* stream);
/* dh_hc_sf_and_derivs_dev with the multipole-parameter derivatives: selector may add
 * DH_D_MULTIPOLE, and d_plm (25 n_types), d_p_val, d_kappa, d_kappa_prime (n_types each) device
 * doubles (any may be NULL) receive dT/dP_lm (local frame, [l*l + l + m]), dT/dPval, dT/dkappa and
 * dT/dkappa' summed over the atoms of each type, the values dh_hc_fetch_type_derivatives returns
 * after the host call (to rounding: the sums run in another order), with P_lm entries beyond a
 * type's multipole order 0.  Without DH_D_MULTIPOLE it is dh_hc_sf_and_derivs_dev and leaves the
 * four arrays unwritten.  Replaces the host call with DH_D_MULTIPOLE (dh_hc_sf_and_derivs +
 * dh_hc_fetch_type_derivatives) for callers whose parameters live on the GPU (torch autograd);
 * the reference has no device-memory form and no multipole-parameter derivatives in
 * calculateStructureFactorsAndDerivatives (.h:148-165). */
int dh_hc_sf_and_derivs_mp_dev(dh_calculator* calc, const dh_atoms_dev* atoms, const double* dtarget_df,
                               unsigned selector, double* f, double* d_xyz, double* d_adp, double* d_occ,
                               double* d_plm, double* d_p_val, double* d_kappa, double* d_kappa_prime, void* stream);
/* dh_hc_sf_and_derivs_dev without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid allocation).  Same arguments, results and stream ordering; the call makes
 * no hipMalloc / hipFree / pinned or library allocation.  Its per-call device buffers (atom
 * records, sorted dT/dF, per-wave F and per-chunk derivative partial sums, sorted F, raw
 * derivatives) come from one static scratch array of DH_NOALLOC_SCRATCH_BYTES per GPU, which every
 * calculator on that GPU shares and which keeps the last call's contents: allocation-free calls
 * that overlap on one GPU (two host threads, or two streams without ordering between them) race.
 * Everything a call reads from the scratch it writes first.  The reflection records come from
 * dh_hc_set_reflections, which must have run in the current calculation mode (and electron
 * setting).  A call that needs more than the scratch (about 5000 atoms x 66 000 reflections in
 * FP64; the reflection chunk is lengthened to fit the derivative partials) or records not
 * prepared returns DH_E_MEMORY without computing anything. */
#define DH_NOALLOC_SCRATCH_BYTES (128u << 20)
int dh_hc_sf_and_derivs_dev_noalloc(dh_calculator* calc, const dh_atoms_dev* atoms, const double* dtarget_df,
                                    unsigned selector, double* f, double* d_xyz, double* d_adp, double* d_occ,
                                    void* stream);

/* ---- Per-reflection derivatives: the reference's per-hkl overload
 *   template<T> calculateStructureFactorsAndDerivatives(atoms, lcs, hkl, T& singleHklResultsProcessor
 *       [, countAtomContribution])     HansenCoppensStructureFactorCalculator.h:206-242, 384-545
 * which runs HansenCoppens_SF_Engine::calculateSF<T> (HansenCoppens_SF_Engine.h:512-780) and calls
 *   singleHklResultsProcessor.onPerHklCalculation(hkl_idx, F, SfDerivativesAtHkl)
 * once per reflection in input order.  Each reflection's derivatives form one row of
 * DH_JAC_PLANES planes of n_atoms complex values (re, im interleaved), in the calculator's
 * derivative conventions (as the reference's WrapperForSingleHklResultProcessor converts them):
 *   plane 0-2  dF/dx, dF/dy, dF/dz          (SfDerivativesAtHkl::atomicPostionDerivatives)
 *   plane 3-8  dF/dU11 .. dF/dU23           (adpDerivatives; U_iso atoms: plane 3 only, 4-8 zero;
 *                                            atoms without ADPs: zero)
 *   plane 9    dF/docc                      (occupancyDerivatives)
 * F follows the per-hkl engine, which differs from the summed one for centrosymmetric groups
 * with f'' != 0: it keeps f'' unconjugated for the inversion partner.  Atoms excluded by
 * countAtomContribution (dh_atoms.include) have zero entries and no part in F.  The rows are
 * FP64; their terms are FP64 in the FP64 modes and FP32 with an FP64 phase reduction in
 * HIP_MIXED / HIP_IAM_MIXED (1e-6 / 1e-5 of the FP64 rows). */
#define DH_JAC_PLANES 10
/* Called once per reflection, hkl_index = 0 .. n_hkl-1 in order; f = 2 doubles, row =
 * 2 * DH_JAC_PLANES * n_atoms doubles, valid during the call only.  Return 0 to continue; any
 * other value stops the calculation, which then returns DH_E_CALLBACK. */
typedef int (*dh_per_hkl_fn)(void* user, int hkl_index, const double* f, const double* row);
int dh_hc_per_hkl(dh_calculator* calc, const dh_atoms* atoms, int n_hkl, const int* hkl, dh_per_hkl_fn fn,
                  void* user);
/* The same rows in bulk: upload once, then fetch any reflection range [k0, k1) into host memory
 * (f: 2 (k1-k0) doubles, rows: (k1-k0) rows; either may be NULL) ... */
int dh_hc_jacobian_upload(dh_calculator* calc, const dh_atoms* atoms, int n_hkl, const int* hkl);
int dh_hc_jacobian_rows(dh_calculator* calc, int k0, int k1, double* f, double* rows);
/* ... or into device memory (HIP device pointers, same layout), ordered after the work already
 * queued on `stream` (a hipStream_t; NULL = the null stream), which waits for the rows before its
 * later work; returns without waiting on the host. */
int dh_hc_jacobian_rows_dev(dh_calculator* calc, int k0, int k1, void* f_dev, void* rows_dev, void* stream);

/* ---- Form-factor route: atomic form factors of every atom for a list of reflections,
 *   HansenCoppensStructureFactorCalculator::calculateFormFactors(lcs, const Vector3i& hkl, f, countAtom)
 *   ::calculateFormFactorsCart(lcs, const Vector3d& h_cart, f, countAtom)
 *       (HansenCoppensStructureFactorCalculator.h:171-181, .cpp:1274-1333; engine
 *        HansenCoppens_SF_Engine::calculateFormFactors, HansenCoppens_SF_Engine.cpp:383-495)
 * which the reference computes one reflection per call.  f = f_core(|h|) + Pval f_val(|h|/kappa)
 * (cubic splines on a 0.001 A^-1 grid, as the reference) + the deformation-valence term in the
 * atom's local frame; no anomalous term; excluded atoms 0.  lcs: 9 n_atoms (NULL = identity
 * frames), include: n_atoms or NULL (all).  Output f: n_h * n_atoms complex values (re, im),
 * reflection-major, atoms in model order.  Synchronous, host pointers, any calculation mode. */
int dh_hc_form_factors(dh_calculator* calc, const double* lcs, const unsigned char* include, int n_h, const int* hkl,
                       double* f);
int dh_hc_form_factors_cart(dh_calculator* calc, const double* lcs, const unsigned char* include, int n_h,
                            const double* h_cart, double* f);

/* ---- Local coordinate systems recomputed from the atomic positions (the TAAM route:
 * LocalCoordinateSystem<AtomInCrystalID> definitions, LocalCoordinateSystem.h:20-55, evaluated by
 * LocalCoordinateSystemCalculator::calculate, LocalCoordinateSystemCalculator.cpp:194-238, every
 * step unless frozen, AnyHcCalculator.cpp:317-330).  A reference atom is an atom of the model under
 * a fractional symmetry operation. */
#define DH_LCS_ANY_ORTHOGONAL 0
#define DH_LCS_AVERAGE_DIRECTION 1
#define DH_LCS_AVERAGE_POSITION 2
typedef struct {
    int atom;
    int rot[9];       /* fractional rotation, row-major */
    double trans[3];  /* fractional translation */
} dh_atom_ref;
typedef struct {
    dh_atom_ref central;
    int direction1_type, direction2_type;     /* DH_LCS_* */
    int n_ref1, n_ref2, n_chirality;          /* n_chirality: 0 or 3 */
    const dh_atom_ref* ref1;
    const dh_atom_ref* ref2;
    const dh_atom_ref* chirality;
    int coordinate_1, coordinate_2;           /* 0 X, 1 Y, 2 Z: axis along direction 1, axis in the plane */
    int is_r;                                 /* right-handed */
} dh_lcs_def;
/* The frames of n definitions for atoms at fractional positions xyz (3 per atom): lcs 9n, row-major
 * 3x3 with the local axes as columns (the layout dh_atoms.lcs takes). */
int dh_lcs_compute(const double cell[6], int n_atoms, const double* xyz_frac, int n, const dh_lcs_def* defs,
                   double* lcs);
/* Attach one definition per atom (n = number of atoms; n = 0 detaches): every following upload
 * computes the frames from the uploaded positions and ignores dh_atoms.lcs.  chain_rule != 0 adds
 * dT/dL : dL/dxyz to the xyz derivatives (the reference treats frames as constant,
 * HansenCoppens_SF_Engine3.cpp:1030-1509); it runs the multipole-adjoint kernels for that. */
int dh_hc_set_lcs_definitions(dh_calculator* calc, int n, const dh_lcs_def* defs, int chain_rule);
/* the frames used by the last upload (9 per atom) */
int dh_hc_last_lcs(dh_calculator* calc, double* lcs);

/* ---- Batches of structures (no reference counterpart: the reference computes one structure per
 * call; a refinement of many small structures calls it once per structure).
 * A batch borrows calculators created on one device; they must outlive it.  Inputs are uploaded
 * per calculator with dh_hc_upload, dh_batch_run computes every structure in one launch sequence
 * (per-structure prep, one persistent F + adjoint kernel over all structures, stage-2
 * reductions), and results are read per calculator with dh_hc_fetch.  The precision is the
 * calculators' calculation mode, which must agree.  selector: DH_D_XYZ | DH_D_ADP | DH_D_OCC bits
 * (DH_D_MULTIPOLE is not batched).  use_graph != 0 captures the step into a hipGraph on first
 * use and replays it while every structure keeps its buffer sizes. */
typedef struct dh_batch dh_batch;
int dh_batch_create(dh_calculator* const* calcs, int n, dh_batch** out);
void dh_batch_destroy(dh_batch* batch);
int dh_batch_run(dh_batch* batch, unsigned selector, int use_graph);
/* time the whole batched step over reflection-chunk candidates; keeps and returns the fastest */
int dh_batch_autotune(dh_batch* batch, unsigned selector, int* chunk);
/* reflections per work unit for every structure (0 = autotuned or default) */
int dh_batch_set_tuning(dh_batch* batch, int hkl_chunk);
/* device time of the last dh_batch_run (events around the step), its chunk, work units and
 * persistent grid size */
int dh_batch_last_timing(dh_batch* batch, double* step_ms, int* chunk, int* units, int* grid);

#ifdef __cplusplus
}
#endif
#endif
