/*
 * gain_engine — MI355X (gfx950) line-strength engine, C ABI.
 *
 * Replaces the device layer of GAIN-MPI's dipole path: the kernels of
 * GPU/dipole_GPU.cu (expand, K-block half line strength), the cuBLAS ddot
 * projection, and the S -> Einstein-A -> intensity stage of
 * BaseClasses/Output.cpp, behind plain C entry points (no torch types).
 *
 * Conventions
 *   - every function returns a gain_status (0 = OK); on failure
 *     gain_last_error() returns a message for the calling thread.  Nothing
 *     aborts the process (the reference calls MPI_Abort,
 *     GpuManager.cpp:80-93, BaseProcess.cpp:42-52).
 *   - host arrays are caller-owned and only read (or written) during the
 *     call; device arrays created by the engine are engine-owned.  Arrays
 *     marked "device" are caller-owned device memory the engine borrows.
 *   - a context is bound to one GPU and one HIP stream; calls on one context
 *     must not overlap (the reference's per-`proc` slots map to separate
 *     contexts).
 */
#ifndef GAIN_ENGINE_H
#define GAIN_ENGINE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GAIN_OK = 0,
  GAIN_ERR_ARG = 1,     /* invalid argument / size / index */
  GAIN_ERR_HIP = 2,     /* HIP runtime error */
  GAIN_ERR_STATE = 3,   /* call order: a prerequisite upload is missing */
  GAIN_ERR_NOMEM = 4,   /* device allocation failed */
  GAIN_ERR_CONFIG = 5   /* inconsistent selection rules / gns */
} gain_status;

typedef struct gain_ctx gain_ctx;

#define GAIN_MAX_IRREPS 16

/* Intensity / selection parameters: the Input.h getters consumed by
 * States (States.cpp:7-38) and Output (Output.cpp:3-31). */
typedef struct {
  int nrepres;                      /* number of irreps */
  int maxdeg;                       /* max irrep degeneracy (<= 3) */
  int maxJ;                         /* largest J of the J list */
  int reduced;                      /* SYMMETRY REDUCED */
  int symmetry_type;                /* 1 generic, 2 C3v/D3h, 3 Td */
  int sym_degen[GAIN_MAX_IRREPS];
  int isym_do[GAIN_MAX_IRREPS];
  int isym_pairs[GAIN_MAX_IRREPS];  /* SELECTION */
  double gns[GAIN_MAX_IRREPS];      /* GNS */
  double erange_lower[2];           /* ENERGY LOW a b */
  double erange_upper[2];           /* ENERGY ... UPPER c d */
  double freq_window[2];            /* FREQ-WINDOW */
  double zpe;                       /* ZPE */
  double temperature;               /* TEMPERATURE */
  double partition;                 /* QSTAT */
  double threshold;                 /* THRESH_LINE (applies to A) */
} gain_params;

/* Per-pair results for a tile of initial states x one final block.
 * Row i, column F (F = level index inside the final block, 0..nF-1).
 * Any pointer may be NULL (not written).  `on_device` = 1 means the
 * pointers are device memory (written without a copy). */
typedef struct {
  double* ls;      /* [nI][nF][maxdeg*maxdeg], ls[dI + dF*maxdeg] */
  double* S;       /* [nI][nF] degeneracy-averaged line strength */
  double* A;       /* [nI][nF] Einstein A */
  double* I;       /* [nI][nF] absorption intensity */
  uint8_t* flags;  /* [nI][nF] bit0: passes the transition filters,
                      bit1: line emitted (nu >= 1e-40 and A >= threshold) */
  int on_device;
} gain_lines;

/* Precision of the contractions of the batched engine. */
typedef enum {
  GAIN_PREC_FP64 = 0,       /* FP64 throughout */
  GAIN_PREC_FP32 = 1,       /* FP32 MFMA products, FP64 accumulation */
  GAIN_PREC_BF16X3 = 2,     /* split-BF16, 3 slices (24 bits of the row max) */
  GAIN_PREC_BF16X4 = 3,     /* split-BF16, 4 slices (32 bits of the row max) */
  GAIN_PREC_I8X5 = 4,       /* split-INT8, 5 slices (35 bits of the row max),
                               exact int32 products on the integer MFMA */
  GAIN_PREC_I8X6 = 5,       /* split-INT8, 6 slices (42 bits) */
  GAIN_PREC_I8X7 = 6,       /* split-INT8, 7 slices (49 bits) */
  GAIN_PREC_I8CRT11 = 7,    /* int8 residues modulo 11 coprime moduli
                               (rows rounded to 43-bit integers relative
                               to their 2-norm), exact integer products by
                               CRT reconstruction */
  GAIN_PREC_I8CRT12 = 8,    /* 12 moduli (47 bits of the row norm) */
  GAIN_PREC_I8CRT14 = 9     /* 14 moduli (47-49 bits of the row maximum) */
} gain_precision;

const char* gain_last_error(void);
const char* gain_version(void);

/* ---- setup (GpuManager ctor + InitializeAndTransferConstants,
 *      GpuManager.cpp:127-213: device, 3j table, constants) ------------ */
int gain_create(gain_ctx** ctx, int device, int jmax, int nrepres,
                int maxdegen);
int gain_destroy(gain_ctx* ctx);
/* The (jmax+1)^2*9 3j table (GpuManager.cpp:164-184 index layout). */
int gain_get_threej(gain_ctx* ctx, double* host_out);
/* HIP stream the context launches on (hipStream_t as void*). */
int gain_get_stream(gain_ctx* ctx, void** stream);
int gain_synchronize(gain_ctx* ctx);

/* Selection rules and intensity constants; computes the irrep pairs like
 * TroveInput::ComputeIGammaPair (TroveInput.cpp:202-236) and returns
 * GAIN_ERR_CONFIG where the reference aborts. */
int gain_set_params(gain_ctx* ctx, const gain_params* p);
int gain_get_gamma_pair(gain_ctx* ctx, int gamma, int* pair);

/* ---- data uploads ------------------------------------------------------ */
/* Dipole mu in the reference piece layout (TroveDipole.cpp:110):
 * mu[(vF-startF) + vI*(endF-startF) + alpha*(endF-startF)*nc];
 * GpuManager::TransferDipole, GpuManager.cpp:499-527.  The engine keeps the
 * whole matrix resident: startF=0, endF=nc replaces it; a piece
 * [startF, endF) (the reference's split of a dipole that does not fit,
 * TroveDipole.cpp:47-112) is copied into the resident matrix, and every
 * product returns GAIN_ERR_STATE until the pieces cover all nc final
 * columns (a piece of another nc starts a new matrix). */
int gain_upload_dipole(gain_ctx* ctx, const double* mu, int nc, int startF,
                       int endF);
/* Primitive basis of one J (GpuManager::TransferBasisSet,
 * GpuManager.cpp:215-265): K-block starts/sizes (J+1 each), tau (= ktau&1)
 * and vibrational index per primitive, primitives K-ordered. */
int gain_upload_basis(gain_ctx* ctx, int jind, int J, int nprim,
                      const int* kstart, const int* kblock, const int* ktau,
                      const int* vib);
/* Symmetric-top (ROTSYM) basis of one J, the basis TroveBasisSet builds
 * with do_rotsym (TroveBasisSet.cpp:480-483: per primitive the rotational
 * level, stored where the K-block basis keeps ktau, and the degenerate
 * component, stored where it keeps K; no K blocks), for the half line
 * strength of dipole_GPU.cu:641-723 (GpuManager::ExecuteRotSymHalfLs,
 * GpuManager.cpp:547-600, chosen by rotsym_do at :612-618).  `nlevels`
 * symmetrised rotational levels of up to `maxdeg` components (the layout
 * dimensions of the rotational table).  Both bases of a pair must be of
 * the same kind; the adjoint covers K-block bases only. */
int gain_upload_basis_rotsym(gain_ctx* ctx, int jind, int J, int nprim,
                             int nlevels, int maxdeg, const int* level,
                             const int* deg, const int* vib);
/* Rotational (Wigner) factors of the symmetric-top pair J index jindI ->
 * jindF, the table GpuManager::TransferWigner uploads
 * (GpuManager.cpp:440-470) and c_wignerGen fills (TroveBasisSet.cpp:
 * 611-633): rot[3 x nlevelsI x nlevelsF x maxdeg x maxdeg], first index
 * fastest (gain_rotsym_wigner produces it for a group). */
int gain_upload_wigner(gain_ctx* ctx, int jindI, int jindF, int nlevelsI,
                       int nlevelsF, int maxdeg, const double* rot);
/* Inflation data of one J (GpuManager::TransferInflation,
 * GpuManager.cpp:381-422): ijterms[maxsymcoeffs*nrepres], icontr[2*nprim],
 * and per irrep N[maxsymcoeffs], repres[ntot*sdeg*matsize]. */
int gain_upload_inflation(gain_ctx* ctx, int jind, int maxsymcoeffs,
                          const int* ijterms, const int* icontr, int matsize,
                          const int* ntot, const int* sdeg, const int** N,
                          const double** repres);
/* All states of one (J index, irrep) block: energies (host, one per state)
 * and the eigenvector records (one record of rec_len doubles per state,
 * EigenVector.cpp:152-178), copied to HBM where they stay resident.
 * `vectors` may be NULL to register energies only (initial-state blocks
 * whose vectors are passed per call). */
int gain_upload_block(gain_ctx* ctx, int jind, int gamma, int nstates,
                      int rec_len, const double* energies,
                      const double* vectors);
/* Same, but the records already live in device memory (row stride `ld`
 * doubles); the engine borrows the pointer.  The split precision modes keep
 * a sliced copy of a block's records, built on first use: after changing
 * the records in place, bind the block again (which drops that copy). */
int gain_bind_block_device(gain_ctx* ctx, int jind, int gamma, int nstates,
                           int rec_len, int64_t ld, const double* energies,
                           const double* vectors_device);

/* ---- hot path ---------------------------------------------------------- */
/* Half line strength of one initial vector, J_I -> J_F, irrep gammaI,
 * component idegI: TransformHalfLsVector + ExecuteKBlockHalfLs +
 * GetHalfLineStrengthResult (GpuManager.cpp:612-752).  cI is host memory
 * of length ntot(jindI, gammaI); h_out receives nprim(jindF) doubles. */
int gain_half_ls(gain_ctx* ctx, int jindI, int gammaI, int idegI,
                 const double* cI, int jindF, double* h_out);

/* Plain FP64 drop-in for one initial state: half line strength, projection
 * of every final state of block (jindF, gammaF) by FP64 GEMV over the
 * HBM-resident records, and the S/A/I epilogue with the transition
 * filters (main.cpp:238-338, GpuManager.cpp:733-759, Output.cpp:52-112).
 * The initial state is level `levelI` of block (jindI, gammaI); cI is its
 * record in host memory, or NULL to read it from the uploaded block. */
int gain_plain_lines(gain_ctx* ctx, int jindI, int gammaI, int levelI,
                     const double* cI, int jindF, int gammaF,
                     const gain_lines* out);

/* Allocation-free form of gain_plain_lines.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same launches, arguments and results as gain_plain_lines, but
 * every work buffer lives in one fixed device array of
 * GAIN_NOALLOC_SCRATCH_BYTES that each GPU holds from the first context on
 * it to process end, and the call makes no device or pinned-host
 * allocation, memset or host-to-device copy.  cI_device: the initial record
 * in device memory, or NULL for the uploaded block's record of levelI.
 * `out` must be device memory (on_device = 1).  A call whose buffers exceed
 * the array returns GAIN_ERR_ARG (the tuning keys "mu_ksplit" and
 * "gemv_ksplit" shrink them).  Returns once queued on the context's stream
 * (gain_synchronize waits).  Every call on one GPU shares that GPU's array,
 * which keeps the last call's contents: calls that overlap on one GPU
 * (several contexts, threads or streams) race. */
#define GAIN_NOALLOC_SCRATCH_BYTES (64ull << 20)
int gain_plain_lines_noalloc(gain_ctx* ctx, int jindI, int gammaI,
                             int levelI, const double* cI_device, int jindF,
                             int gammaF, const gain_lines* out);

/* Batched engine: nI initial states of block (jindI, gammaI) (levels
 * levelsI[0..nI), vectors from the uploaded block) against every final
 * state of (jindF, gammaF); half line strengths and projection run as
 * GEMMs at the requested precision.  Rows of `out` follow levelsI. */
int gain_batched_lines(gain_ctx* ctx, int jindI, int gammaI,
                       const int* levelsI, int nI, int jindF, int gammaF,
                       int precision, const gain_lines* out);

/* ---- mixed-J sweeps (one HIP graph per sweep) ---------------------------- *
 * A sweep is a fixed list of tasks, each an initial batch of one
 * (J index, irrep) block against one final block (the work of many
 * gain_batched_lines calls, e.g. every J_I -> J_I-1, J_I, J_I+1 pair of a J
 * list).  At creation the engine plans every task, runs it once, and
 * captures the launch sequence in a HIP graph: the initial batches
 * (tasks sharing an initial batch share their half line strengths) are
 * ordered by J and cut into phases (tuning key "sweep_phases"); the mu X
 * products of a phase run as one GEMM on the context's stream, and
 * everything downstream of it (gathers, projections as one persistent
 * grouped FP64 MFMA launch and one grouped GEMV launch, every per-task
 * kernel as one grouped launch per stage) on a second stream of the sweep,
 * beside the next phase's products.  `precision` is GAIN_PREC_FP64,
 * GAIN_PREC_FP32, a split-BF16 mode (GAIN_PREC_BF16X3/4), a split-INT8 mode
 * (GAIN_PREC_I8X5..7) or a CRT mode (GAIN_PREC_I8CRT11/12/14); the reduced
 * modes apply to the mu X products only (the projections stay FP64).
 * gain_sweep_run replays the graph on the context's stream
 * (asynchronously; gain_synchronize waits).  Task outputs must be device
 * memory.  Re-uploading the dipole, a basis or a block used by a sweep
 * invalidates it.  No counterpart in the reference (its loop is
 * main.cpp:238-338). */
typedef struct {
  int jindI, gammaI;
  const int* levelsI;  /* host, nI levels of block (jindI, gammaI) */
  int nI;
  int jindF, gammaF;
  gain_lines out;      /* [nI][nF] results, device memory */
} gain_task;
typedef struct gain_sweep gain_sweep;
int gain_sweep_create(gain_ctx* ctx, const gain_task* tasks, int ntasks,
                      int precision, gain_sweep** sweep);
int gain_sweep_run(gain_sweep* sweep);
int gain_sweep_destroy(gain_sweep* sweep);

/* ---- adjoint ------------------------------------------------------------- *
 * No counterpart in the reference (it has no derivatives).  Gradient of a
 * loss L with respect to the dipole, given dS = dL/dS for the pairs of a
 * batch (same layout as gain_lines.S): recomputes the FP64 forward, then
 * dls = dS dS/dls, dG = C_F^T dls, dH = P_F dG, dY = (gather)^T dH and
 * dL/dmu_alpha = dY_alpha X^T (a rank-k update).  dmu has the dipole layout
 * of gain_upload_dipole; `accumulate` adds into it.  dS and dmu are device
 * memory if on_device, else host memory. */
int gain_adjoint_dipole(gain_ctx* ctx, int jindI, int gammaI,
                        const int* levelsI, int nI, int jindF, int gammaF,
                        const double* dS, double* dmu, int on_device,
                        int accumulate);
/* The same without recomputing the forward: ls and flags are the
 * gain_lines.ls and gain_lines.flags that the FP64 batched forward
 * (gain_batched_lines, GAIN_PREC_FP64) wrote for this batch and block with
 * the current dipole ([nI][nF][maxdeg*maxdeg] and [nI][nF]; device memory
 * if on_device, else host).  Only the scattered initial vectors of the
 * rank-k update are rebuilt.  The caller keeps them matched: results for
 * another batch, dipole or block give a wrong gradient. */
int gain_adjoint_dipole_ls(gain_ctx* ctx, int jindI, int gammaI,
                           const int* levelsI, int nI, int jindF, int gammaF,
                           const double* dS, const double* ls,
                           const unsigned char* flags, double* dmu,
                           int on_device, int accumulate);
/* mu = sum_{p<P} coef[p] M_p (coef host, M device [P][3][nc][nc]); sets the
 * context's dipole. */
int gain_set_dipole_combination(gain_ctx* ctx, int P, const double* coef,
                                const double* M_device, int nc);
/* dc[p] = <M_p, dmu> over all 3 nc^2 entries (M, dmu device; dc host). */
int gain_dipole_coeff_grad(gain_ctx* ctx, int P, const double* M_device,
                           const double* dmu_device, double* dc);

/* ---- per-kernel timing (HIP events around each launch) ------------------ */
int gain_set_timing(gain_ctx* ctx, int enable);
/* Accumulated time (ms) and launch count of a named kernel family:
 * "gemv", "gemm_proj" (batched projection stage), "proj_gemm" (its matrix
 * product kernel alone), "half_ls" (half line strength stage),
 * "half_ls_gemm" (its mu X product kernel alone), "prep" (one-time builds
 * of the reduced-precision planes of the dipole and of final blocks),
 * "expand_T", "epilogue", "adjoint" (the
 * backward of gain_adjoint_dipole after its forward recomputation) and
 * "adjoint_rankk" (its rank-k update dL/dmu = dY X^T). Resets the counters
 * if reset != 0. */
int gain_get_timing(gain_ctx* ctx, const char* name, double* total_ms,
                    int64_t* launches, int reset);

/* ---- tuning -------------------------------------------------------------- *
 * Keys: "autotune" (1: time every GEMM tile configuration on the first call
 * of each shape and keep the fastest; 2: forget earlier choices),
 * "proj_cfg" / "half_ls_cfg" (force a configuration, -1 = automatic),
 * "i8_cfg" / "i8_split" (split-INT8 GEMMs: force a tile configuration
 * 0..11, -1 = automatic; projection split count, 0 = automatic),
 * "crt_cfg" (CRT-form int8 GEMMs: force a tile configuration 0..13),
 * "crt_kskip" (1, the default: the CRT-form dipole products multiply only
 * the k blocks of the dipole's residue planes that hold a nonzero entry,
 * with results identical to the dense products; 0: every k block),
 * "proj_crt_moduli" (0, the default: the CRT-form projection uses the
 * precision's moduli count; 11, 12 or 14: its own, the half line strength
 * keeping the precision's),
 * "f32_cfg" (FP32 matrix-core GEMMs: force a tile configuration 0..3,
 * -1 = the default, 0),
 * "crt_fuse" (1, the default: the CRT-form half line strength of K-block
 * bases is reconstructed and gathered by one kernel, identical results
 * without the FP64 mu X intermediate; 0: the two kernels),
 * "crt_xfuse" (1, the default: the CRT-form half line strength builds the
 * initial vectors' X straight into its residue planes, one kernel, when
 * N_c <= 4096; 0: the FP64 X scattered, then cut by two kernels; the
 * residues are identical),
 * "gemv_rows" (2/4/8), "gemv_ksplit", "mu_ksplit" (0 = automatic),
 * "mfma" (0: vector-ALU GEMMs, for cross-checks), "noalloc_cap" (bytes of
 * the allocation-free entry point's array a call may use, 0 = all),
 * "sweep_phases" (1..16, default 2: pipeline phases of sweeps created
 * afterwards; 1 = one stream), "sweep_eager" (1: gain_sweep_run launches
 * the sweep's kernels one by one instead of replaying its graph). */
int gain_set_tuning(gain_ctx* ctx, const char* key, int value);
/* Current value of a settable key, or what the most recent launch used:
 * "proj_cfg", "proj_split", "proj_grid" (FP64 MFMA projection GEMM),
 * "half_ls_cfg", "half_ls_split", "half_ls_grid" (mu GEMM),
 * "half_ls_kblocks_ppm" (CRT form: k blocks the mu GEMM multiplies, parts
 * per million of the dense count), "gemv_grid"
 * (plain GEMV); *_grid = work-items of that launch; "sweep_phases",
 * "sweep_gemv_packs", "sweep_gemv_bytes" (the latest sweep created);
 * "autotune_us" = host
 * time spent timing tile configurations so far. No reference
 * counterpart: the reference has no tunable kernels. */
int gain_get_tuning(gain_ctx* ctx, const char* key, long long* value);
/* The autotuner's timings and choices, one line per GEMM shape. */
int gain_tuning_report(gain_ctx* ctx, char* buf, int len);
/* Measured matrix-core ceiling: register-only chains of one MFMA
 * instruction ("i8": v_mfma_i32_32x32x32_i8, "f64": v_mfma_f64_16x16x4f64)
 * on every CU, in operations (2 per multiply-add) per second.  No reference
 * counterpart: it calibrates the rooflines of the split-INT8 and FP64
 * GEMMs. */
int gain_mfma_peak(gain_ctx* ctx, const char* kind, double* ops_per_s);
/* Read-only HBM rate of a plain streaming kernel (16-byte non-temporal
 * loads over an 8 GiB buffer allocated for the call), in bytes/s: the
 * achievable ceiling for the HBM-bound kernels. */
int gain_hbm_read_peak(gain_ctx* ctx, double* bytes_per_s);
/* The CRT form's moduli and reconstruction, on the host (no GPU needed).
 * No reference counterpart: they expose the integer arithmetic of the
 * int8 CRT kernels for checks against exact integers.
 * gain_crt_moduli: m[0..nm) for nm = 11, 12 or 14.  gain_crt_beta: the
 * integer width of the operands of a product of length K, and (if
 * norm_scaled is not NULL) 1 when rows are scaled by their 2-norm (the
 * products then bounded by Cauchy-Schwarz), 0 when by their maximum.
 * gain_crt_reconstruct: words[n4][nm] hold the symmetric residues of four
 * outputs (byte q of word i = residue of output q modulo m_i); out[n4][4]
 * receives the integer products in FP64, computed as the GPU kernels do. */
int gain_crt_moduli(int nm, int* m);
int gain_crt_beta(int nm, long long K, int* beta, int* norm_scaled);
int gain_crt_reconstruct(int nm, const unsigned int* words, long long n4,
                         double* out);

/* ---- transition list (main.cpp:238-338 + Output.cpp:52-112) ----------- */
typedef struct {
  int full_line;          /* print the components "[ ... ] " (Output.cpp:64,105) */
  int compute_intensity;  /* print the intensity column (Output.cpp:100-103) */
  int batch;              /* initial states per batched call (0: 64) */
  int precision;          /* gain_precision of the contractions */
} gain_output_opts;
/* One line exactly as Output::OutputLinestrength prints it
 * ("%12.6f %8d %4d %4d <- %8d %4d %4d %16.9E " ... " || \n"); levels and
 * irreps are 0-based and printed +1 as the reference does.  nls = -1: no
 * components block; nls >= 0: "[ " + nls x " %16.9E" + " ] ". */
int gain_format_line(char* buf, int len, double nu, int levelF, int jF,
                     int gammaF, int levelI, int jI, int gammaI, double A,
                     int with_intensity, double intensity, const double* ls,
                     int nls);
/* Every line of every uploaded (J, gamma) block against every block the
 * selection rules allow (the main.cpp:238-338 loop: FilterLowerState,
 * irrep pair, |dJ| <= 1, FilterIntensity, threshold), written to `path`
 * (NULL: stdout).  Initial states in energy order, each one's finals in
 * energy order.  All participating blocks need their eigenvectors
 * resident.  *nlines = lines written. */
int gain_write_transitions(gain_ctx* ctx, const gain_output_opts* opts,
                           const char* path, int64_t* nlines);

/* ---- TROVE checkpoint files (host-side readers) ------------------------ */
/* Level number printed for each state of a block (the descr file's ilevel,
 * 0-based; TroveStates.cpp:94); NULL restores the row index. */
int gain_set_block_levels(gain_ctx* ctx, int jind, int gamma,
                          const int* levels);
/* j0_extfield.chk (TroveDipole.cpp:9-119): Fortran sequential records,
 * including >2 GiB subrecords (Util.cpp:89-109).  Returns nc; with piece
 * != NULL also fills the 3*nc*nc piece layout gain_upload_dipole takes
 * (transposed per component as TroveDipole.cpp:110 does). */
int gain_trove_read_dipole(const char* path, int* nc, double* piece);
/* j0eigen_quanta<J>.chk plus the J=0 file (TroveBasisSet.cpp:27-709):
 * contracted basis, correlation with J=0 (vibrational index per function),
 * K-blocks, tau, icontr, ijterms, N and repres per irrep. */
typedef struct gain_trove_basis gain_trove_basis;
int gain_trove_read_basis(const char* path, const char* path_j0, int J,
                          int nrepres, const int* sdeg,
                          gain_trove_basis** out);
int gain_trove_basis_free(gain_trove_basis* b);
/* ntot has nrepres entries. */
int gain_trove_basis_sizes(const gain_trove_basis* b, int* nprim,
                           int* maxsymcoeffs, int* matsize, int* ntot);
/* Copy one array out: "kstart", "kblock" (J+1 ints), "tau", "vib"
 * (nprim ints), "icontr" (2*nprim), "ijterms" (maxsymcoeffs*nrepres),
 * "N" (maxsymcoeffs ints of irrep gamma), "repres" (doubles of gamma). */
int gain_trove_basis_get(const gain_trove_basis* b, const char* name,
                         int gamma, void* dst);
/* gain_upload_basis + gain_upload_inflation from a read basis. */
int gain_trove_upload_basis(gain_ctx* ctx, int jind,
                            const gain_trove_basis* b);
/* j0eigen_descr<J>_<gamma+1>.chk (TroveStates.cpp:15-131): the states that
 * pass the state filter (lower or upper window, isym_do; States.cpp:41-53)
 * in file order.  Arrays may be NULL (count only): irec and ilevel are
 * 0-based, rec_len = the file's dim_basis. */
int gain_trove_read_states(const char* path, int gamma, const gain_params* p,
                           int* nstates, int* rec_len, int* irec,
                           int* ilevel, double* energy);
/* j0eigen_vectors<J>_<gamma+1>.chk: record irec[i] (rec_len doubles at
 * irec*rec_len, EigenVector.cpp:88) into out[i*rec_len]. */
int gain_trove_read_records(const char* path, int n, const int* irec,
                            int rec_len, double* out);
/* Everything the reference's main.cpp:118-212 reads, from directory `dir`,
 * for J list jvals (J index = position): dipole, bases, and for every irrep
 * with isym_do the states and their records (uploaded, level numbers
 * set).  gain_set_params must come first. */
int gain_trove_load(gain_ctx* ctx, const char* dir, const int* jvals,
                    int nj);

/* ---- kernel-level drop-ins for GPU/dipole_GPU.h ------------------------- *
 * Same argument meaning as the reference's free functions; pointers are
 * device memory, `stream` is a hipStream_t (NULL = default stream).
 * The __constant__ setters mirror copy_*_constant (dipole_GPU.cu:119-134). */
int gain_copy_symmetry_constants(int sym_nrepres, int maxdeg);
int gain_copy_jmax_constant(int jmax);
int gain_copy_dipole_constant(int dip_stride);
/* transform_vector_primitive, dipole_GPU.h:19 */
int gain_transform_vector_primitive(int dimenI, int igammaI, int maxcoeff,
                                    int idegI, int sdeg, int Ntot,
                                    const int* ijterms, const int* icontr,
                                    const int* N, const double* repres,
                                    const double* vecI, double* vec,
                                    void* stream);
/* compute_gpu_half_linestrength_, dipole_GPU.h:23-26 (one K_F block) */
int gain_compute_half_linestrength(
    int dimenF, int dimenI, int jI, int jF, int kF, const int* tauI,
    const int* tauF, const int* icorrI, const int* icorrF, int startF_idx,
    int startI_idx, int startFblock, int endFblock, int ncontrF,
    int kFBlocksize, const int* kblock_size, const double* dipole_me,
    const double* vector, const double* threej, double* half_ls,
    void* stream);
/* compute_gpu_half_linestrength_rotsym_, dipole_GPU.h:38-41 (kernel
 * dipole_GPU.cu:641-723): half line strength in the symmetric-top basis.
 * kI/kF = degeneracy index, ktauI/ktauF = rotational level, icorrI/icorrF =
 * vibrational index of each primitive; wigner holds
 * 3 x nlevelI x nlevelF x sym_max_degen x sym_max_degen doubles
 * (first index fastest).  Needs gain_copy_dipole_constant. */
int gain_compute_half_linestrength_rotsym(
    int dimenF, int dimenI, int nlevelI, int nlevelF, int sym_max_degen,
    const int* kI, const int* kF, const int* ktauI, const int* ktauF,
    const int* icorrI, const int* icorrF, int startFblock, int endFblock,
    int ncontrF, const double* dipole_me, const double* wigner,
    const double* vector, double* half_ls, void* stream);
/* ---- symmetric-top (ROTSYM) rotational factors ---------------------------
 * Groups: "C(M)", "CS(M)", "C2V(M)", "C2H(M)", "CS(EM)", "G4(M)",
 * "G4(EM)", "D2H(M)", "D2H", "C3V(M)", "D3H(M)", "TD(M)" (symmetry.f90:58-745;
 * the short names "C", "CS", "C2V", "C2H", "G4", "C3V", "D3H", "TD" too)
 * and "DNH(M)" / "DNH" at the order the reference runs it (its order
 * sym%N is never assigned and stays 1: E, sigma_v, i, C'2, all with the
 * identity as equivalent rotation; symmetry.f90:746-1006).  "DINFTYH(M)"
 * is refused: its table (symmetry.f90:1444-1543) leaves the degenerate
 * irreps' matrices and the equivalent rotations of sigma_v and C'_inf
 * unset.
 *
 * Symmetrised rotational functions of one J (MLrotsymmetry_generate,
 * trove_wrappers.f90:416-589, with MLsymmetrization_rot :591-914): the
 * 2J+1 Wang functions (k = 0, then k = 1..J with tau = 0, 1) combined into
 * nlevels levels, each of one irrep (level_irrep[nlevels]) and with
 * maxdeg components; root_of[level * maxdeg + deg] is the column of the
 * function in vectors[(2J+1) x (2J+1)] (row = Wang function, row major),
 * -1 past the irrep's degeneracy.  operations (nullable) receives the real
 * part of the rotation matrix of every group operation in the Wang basis,
 * noper x (2J+1) x (2J+1).  Output arrays may be NULL (sizes only). */
int gain_rotsym_functions(const char* group, int J, int* nlevels,
                          int* maxdeg, int* level_irrep, int* root_of,
                          double* vectors, double* operations);
/* c_wignerGen (contraced_rotational_dipole, trove_wrappers.f90:170-414):
 * the rotational dipole factors between the symmetrised functions of J_I
 * and J_F (|J_F - J_I| <= 1), rot[3 x nlevelsI x nlevelsF x maxdeg x maxdeg]
 * (first index fastest), the table gain_compute_half_linestrength_rotsym
 * takes.  rot may be NULL (sizes only). */
int gain_rotsym_wigner(const char* group, int jI, int jF, int* nlevelsI,
                       int* nlevelsF, int* maxdeg, double* rot);
/* cublasDdot(handle, n, x, 1, y, 1, result) in DEVICE pointer mode,
 * GpuManager.cpp:741 */
int gain_ddot(int n, const double* x, const double* y, double* result,
              void* stream);

#ifdef __cplusplus
}
#endif
#endif
