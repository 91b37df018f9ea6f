/*
 * su2dmrg.h -- C ABI of the MI355X SU(2) DMRG hot path.
 *
 * Plain pointers and sizes only.  Every call returns an int status
 * (SU2_OK = 0; see su2_status below) where the reference prints and calls
 * exit(0) (e.g. tensor_su2.cpp:165-168, tensor_su2_dmrg_src.cpp:40-53);
 * su2_last_error() returns the message of the last failure on a context.
 *
 * Conventions shared with the reference (lingwangqs/SU2DMRG):
 *  - spins are doubled: twoj = 2j (tensor.cpp:1355);
 *  - a leg is a list of sectors (twoj ascending, multiplet count) plus a
 *    direction +-1 (su2bond.hpp:3-29);
 *  - blocks are indexed mixed-radix over sector indices with leg 0 fastest,
 *    and a block exists when sequential fusion of its labels reaches the
 *    tensor's locspin (su2struct.cpp:243-332);
 *  - block data are the reduced tensors, column-major with leg 0 fastest
 *    (tensor.cpp:1101-1119), concatenated in block-index order -- the
 *    serialisation of tensor_su2::get_telement(tele1,tele2,iflag)
 *    (tensor_su2.cpp:532-549) without the CG part, which is implicit:
 *    two-leg wavefunctions carry the unit singlet, two-leg operators the
 *    identity, three-leg operators make_cgc(j0,j1,j2).
 *
 * Calls on one context are serialised on its stream and must not overlap
 * from several host threads (the reference calls these kernels from inside
 * OpenMP regions, dmrg_su2.cpp:582 -> tensor_su2_dmrg_src.cpp:977; here the
 * caller hands the whole term list to one call instead).
 */
#ifndef SU2DMRG_H
#define SU2DMRG_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct su2_ctx su2_ctx;
typedef struct su2_tensor su2_tensor;
typedef struct su2_heff su2_heff;
typedef struct su2_dmrg su2_dmrg;
typedef struct su2_batch su2_batch;
typedef struct su2_server su2_server;

/* One served job (the flags of one reference run, main_omp.cpp:42-132:
 * -lx -ly -jcoup -qcoup -boundary -sec -d -exci and the sweep count). */
typedef struct su2_job {
  int lx, ly;
  double jcoup, qcoup;
  int bdry, totspin, max_dcut;
  int exci;      /* 0: ground state; k > 0: k-th excited state of the sector */
  int nsweep;    /* sweeps after the warm-up, per level */
  double stop;   /* |dE| early stop between sweeps (main_omp.cpp:197); < 0: off */
} su2_job;

enum su2_status {
  SU2_OK = 0,
  SU2_ERR_ARGUMENT = 1,
  SU2_ERR_STRUCTURE = 2,
  SU2_ERR_DEVICE = 3,
  SU2_ERR_CAPACITY = 4,
  SU2_ERR_NUMERIC = 5
};

/* Library version string. */
const char* su2_version(void);

/* Context on one HIP device (replaces the globals psize/max_dcut/fac_* that
 * tensor_su2.cpp:14-16 and tensor.cpp:25 read). */
int su2_ctx_create(int device, su2_ctx** out);
int su2_ctx_destroy(su2_ctx* ctx);
const char* su2_last_error(const su2_ctx* ctx);
int su2_ctx_synchronize(su2_ctx* ctx);
/* Device memory of the context's cache: bytes in use, the most in use at
 * once since creation (or the last reset_peak = 1), bytes cached for reuse. */
int su2_ctx_memory(su2_ctx* ctx, long long* live, long long* peak, long long* cached,
                   int reset_peak);
/* Tuning of the one-sided Jacobi SVD behind su2_svd_truncate and the DMRG
 * solves: column block width 16 or 32 (default 32), inner Jacobi sweeps per
 * visit of a pair 1..3 (default 1).  No counterpart in the reference; the
 * results agree to the SVD's own accuracy either way. */
int su2_ctx_set_svd_tuning(su2_ctx* ctx, int block_cols, int inner_sweeps);
/* QR preconditioning of the Jacobi SVD: for every sector block with
 * min_k <= min(m, n) <= 2048 (min_k 2..2048, default 256; 0 turns it off),
 * the columns sorted by norm, the Gram matrix's Cholesky factor Rc and the
 * Householder QR Rc^T = Q2 R2 give an orthogonal Q2 with which the Jacobi
 * sweeps start (W <- W P Q2, V <- P Q2).  A block then needs fewer sweeps,
 * and a rank-deficient one (two-site wavefunction sectors typically are)
 * has its null directions in columns at the rounding level from the start,
 * which the sweeps leave out: a 16x8 D=2000 sweep takes about 26 s instead
 * of 37 s.
 * The reference computes its sector SVDs through the Gram matrix and dsyev
 * (tensor.cpp:1472-1529); this setting changes results by rounding only. */
int su2_ctx_set_svd_precondition(su2_ctx* ctx, int min_k);

/* Block-sparse tensor on the device: su2struct + reduced blocks
 * (tensor_su2::set_tensor_su2(su2struct&, double*, double*, int*),
 * tensor_su2.cpp:1431-1470).  twoj/dims hold nsec[0]+..+nsec[nleg-1]
 * entries, leg after leg; present (nullable) holds one flag per block index
 * (iflag of dmrg_su2_omp.cpp:864-941); blocks start zeroed. */
int su2_tensor_create(su2_ctx* ctx, int nleg, const int* nsec, const int* twoj, const int* dims,
                      const int* dirs, int locspin, const int* present, su2_tensor** out);
int su2_tensor_destroy(su2_ctx* ctx, su2_tensor* t);
/* A new zeroed tensor with the legs, locspin and present blocks of src (the
 * structure half of the reference's copy constructor, tensor_su2.cpp:48-58). */
int su2_tensor_like(su2_ctx* ctx, const su2_tensor* src, su2_tensor** out);
/* Number of reduced doubles and of present blocks. */
int su2_tensor_size(const su2_tensor* t, long long* ndouble, int* nblock);
/* Block i (0..nblock-1, block-index order): block index, labels, dims, offset. */
int su2_tensor_block(const su2_tensor* t, int i, int* index, int* twoj, int* dims,
                     long long* offset);
/* Host <-> device copies of all reduced doubles (n must equal the size). */
int su2_tensor_upload(su2_ctx* ctx, su2_tensor* t, const double* host, long long n);
int su2_tensor_download(su2_ctx* ctx, const su2_tensor* t, double* host, long long n);

/* Lanczos vector algebra over the reduced blocks (tensor_su2::operator=,
 * +=, -=, *=, /=, inner_prod, normalize_vector: tensor_su2.cpp:106-259,
 * 324-356, 663-684).  Reductions are deterministic. */
int su2_tensor_copy(su2_ctx* ctx, su2_tensor* dst, const su2_tensor* src);
int su2_tensor_axpy(su2_ctx* ctx, su2_tensor* y, double a, const su2_tensor* x);
int su2_tensor_scal(su2_ctx* ctx, su2_tensor* x, double a);
int su2_tensor_dot(su2_ctx* ctx, const su2_tensor* x, const su2_tensor* y, double* out);
int su2_tensor_normalize(su2_ctx* ctx, su2_tensor* x, double* norm);
/* Allocation-free su2_tensor_dot.  No counterpart in the reference: it is for
 * callers that cannot allocate device or pinned memory (code under HIP graph
 * capture, tight serving loops).  Same kernels and summation order as
 * su2_tensor_dot, hence the same bits; the partial sums go to a fixed
 * per-GPU array of su2_noalloc_scratch_bytes() bytes (4104) that the library
 * holds until the process ends, so the call makes no hipMalloc, hipFree or
 * pinned allocation.  The partial count is fixed, so tensors of every size
 * fit it.  Calls on one GPU share that array (it keeps the last call's
 * contents): calls that overlap on one GPU, from several threads or
 * contexts, race.  The context must have been created (su2_ctx_create
 * resolves the array's address for its device). */
int su2_tensor_dot_noalloc(su2_ctx* ctx, const su2_tensor* x, const su2_tensor* y, double* out);
int su2_noalloc_scratch_bytes(long long* bytes);

/* H_eff plan (dmrg_su2::hamiltonian_vector_multiplication_idmrg,
 * dmrg_su2.cpp:543-664, and tensor_su2::hamiltonian_vector_multiplication,
 * tensor_su2_dmrg_src.cpp:893-1041).
 *   psi_layout : two-leg singlet wavefunction defining the block layout;
 *   hl, hr     : block Hamiltonians hh[il], hh[ir] (nullable); the plan
 *                copies them, so they may be destroyed after this call;
 *   nterm      : number of L . R terms;
 *   term_rank  : doubled rank 2K of each term's operator leg (nullable: all
 *                2, the spin operators of dmrg_su2.cpp:585-610; 0 and 4 are
 *                the rank-0 and rank-2 parts of the J-Q plaquette terms,
 *                dmrg_su2.cpp:635-655), contracted with fac_hamilt_vec[j][2K][j'];
 *   left_count[t] operators of term t are summed as
 *     sum_i left_coef[i] * left_ops[i] (consecutive across terms), likewise
 *     for the right side -- this is the coupling sum op = sum_k coup *
 *     opr[ir][k] of dmrg_su2.cpp:585-599, formed once here.
 * Operators are three-leg [ket, operator leg (2j = term_rank[t], one
 * multiplet), bra]; the plan copies what it needs from them. */
int su2_heff_create(su2_ctx* ctx, const su2_tensor* psi_layout, const su2_tensor* hl,
                    const su2_tensor* hr, int nterm, const int* term_rank,
                    const int* left_count, const su2_tensor* const* left_ops,
                    const double* left_coef, const int* right_count,
                    const su2_tensor* const* right_ops, const double* right_coef,
                    su2_heff** out);
/* out = H_eff psi (psi, out: the plan's layout; out may not alias psi). */
int su2_heff_apply(su2_ctx* ctx, su2_heff* plan, const su2_tensor* psi, su2_tensor* out);
/* Algorithmic FP64 flops of one apply in the reference's GEMM order
 * (per term and block pair: step1 + step2, plus the hh products). */
int su2_heff_flops(const su2_heff* plan, double* algorithmic, double* executed);
int su2_heff_destroy(su2_ctx* ctx, su2_heff* plan);

/* Time `reps` back-to-back applies on the context stream with HIP events
 * (after one untimed apply); returns the mean milliseconds per apply and the
 * mean milliseconds of the two GEMM stages measured separately. */
int su2_heff_time(su2_ctx* ctx, su2_heff* plan, const su2_tensor* psi, su2_tensor* out, int reps,
                  double* ms_apply, double* ms_stage1, double* ms_stage2);

/* Tile-shape autotuner of the two grouped FP64 MFMA launches of a plan: every
 * compiled tile variant is timed (HIP events, `reps` launches) on the plan's
 * own problem list and the fastest is kept.  chosen[2] receives the variant of
 * stage 1 and stage 2; times (nullable) receives 2*nvariants ms per launch
 * (stage 1 first).  *nvariants receives the number of compiled variants.
 * No counterpart in the reference (which calls MKL dgemm per block). */
int su2_heff_autotune(su2_ctx* ctx, su2_heff* plan, int reps, int* chosen, float* times,
                      int* nvariants);
/* Pin the tile variants of the two launches (e.g. to a tuned choice). */
int su2_heff_set_variants(su2_ctx* ctx, su2_heff* plan, int stage1, int stage2);
/* Arithmetic of su2_heff_apply: bits = 64 (default) or 32.  32 runs both
 * grouped launches on FP32 MFMA (v_mfma_f32_16x16x4_f32) with FP32
 * accumulation, on FP32 copies of the plan's operator bundles and block
 * Hamiltonians made on first use; psi is rounded to FP32 and the result
 * widened to FP64 (relative error about 1e-6).  The early sweeps of a
 * mixed-precision run use it (su2_dmrg_set_matvec_precision); the reference
 * computes in FP64 only (tensor_dmrg_src.cpp:536-575).  Tile variants,
 * autotuning and su2_heff_time then apply to the FP32 launches. */
int su2_heff_set_precision(su2_ctx* ctx, su2_heff* plan, int bits);
/* Kernel path of the two launches: uniform = 1 (default) runs a launch whose
 * slices share one operand layout (stage 1: op(A) = A^T, stage 2: op(A) = A)
 * with the straight-line kernel, 0 with the general kernel; both give the
 * same bits.  layouts (nullable) receives the layout of each stage: 0 for
 * op(A) = A, 1 for op(A) = A^T, -1 for the general kernel.  No counterpart in
 * the reference. */
int su2_heff_set_kernel_path(su2_ctx* ctx, su2_heff* plan, int uniform, int* layouts);

/* Truncated SVD of a two-leg singlet wavefunction across sectors
 * (tensor_su2::svd(tu,p1,tv,p2,wout), tensor_su2.cpp:1619-1758, with
 * tensor::svd per sector, tensor.cpp:1472-1700): the singular triplets of all
 * sectors by one batched one-sided block Jacobi SVD on the device (the
 * reference diagonalises Gram matrices and refines their small eigenpairs; the
 * triplets agree to its accuracy), then the reference's keep rule and the
 * global shell-sorted cut to max_dcut.  U = [psi leg 0, new bond] scaled by
 * s^p1, V = [psi leg 1, new bond] scaled by s^p2 (new tensors, destroy with
 * su2_tensor_destroy).  kept (nullable, capacity cap) receives the kept
 * singular values sector by sector, ascending within a sector; *nkept their
 * number; *discarded the summed squares of the cut values. */
int su2_svd_truncate(su2_ctx* ctx, const su2_tensor* psi, double p1, double p2, int max_dcut,
                     su2_tensor** U, su2_tensor** V, double* kept, int cap, int* nkept,
                     double* discarded);

/* One environment kernel on one block operator: the tensor_su2 members
 * that dmrg_su2::prepare_site_operator_from_left/right (dmrg_su2_omp.cpp:
 * 319-825) call, tensor_su2_dmrg_src.cpp:27-890.  The block leg (nsec
 * sectors: twoj ascending, multiplet counts) is enlarged by one spin-1/2 site
 * (side 0: left block, block x site; 1: right block, site x block); the
 * kernel acts there, then the result is projected, O' = P^T O P, when P is
 * given -- P = [enlarged leg, new bond], the U (left) or V (right) that
 * su2_svd_truncate returns for the step (the reference's uu[i]); with P NULL
 * the result stays on the enlarged leg (the reference's uu1, the fusion
 * isometry, dmrg_su2_omp.cpp:604-825).
 *   kind 0 operator_transformation  S_k (x) 1     op: 3-leg spin operator
 *   kind 1 overlap_transformation   H (x) 1       op: 2-leg block Hamiltonian
 *   kind 2 operator_pairup          coef S_k . s  op: 3-leg spin operator,
 *                                                 coef: the coupling J
 *   kind 3 operator_initial         1 (x) s       op: NULL
 * The recoupling factors are the fac_operator_* tables (dmrg_su2.cpp:
 * 759-943) and the site operator carries the reference's sigma = +-sqrt(3)/2
 * (dmrg_su2.cpp:503-506).  *out: a new tensor (3 legs for kinds 0 and 3, 2
 * legs for 1 and 2). */
int su2_env_transform(su2_ctx* ctx, int kind, int side, int nsec, const int* twoj,
                      const int* dims, const su2_tensor* P, const su2_tensor* op, double coef,
                      su2_tensor** out);

/* Recoupling scalars of the environment kernels for one label tuple (doubled
 * spins), the entries of the reference's fac_* tables (dmrg_su2.cpp:759-943):
 *   kind 0 fac_operator_transformation_left  labels (jl, ja, jl', ja', k, s)
 *   kind 1 fac_operator_transformation_rght  labels (jr, ja, jr', ja', k, s)
 *   kind 2 fac_operator_onsite_left          labels (jl, ja, ja', k, s0, s1)
 *   kind 3 fac_operator_onsite_rght          labels (jr, ja, ja', k, s0, s1)
 *   kind 4 fac_operator_pairup_left          labels (ja, jl, jl', k, s0, s1)
 *   kind 5 fac_operator_pairup_rght          labels (ja, jr, jr', k, s0, s1)
 *   kind 6/7 wavefunction re-coupling across one site (jl, s, jr), left->right
 *          and right->left (no reference table: used for start vectors)
 *   kind 8 fac_hamilt_vec                    labels (j, k, j')
 *   kind 9/10  plaquette coupling of a block vector operator and the site
 *          spin to rank K, left / right   labels (jl, ja, jl', ja', 2K, s)
 *   kind 11/12 plaquette reduction of a block rank-K operator and the site
 *          spin to rank 1, left / right   labels (jl, ja, jl', ja', 2K, s)
 *          (the reference's fac_permutation_* role, dmrg_su2.cpp:947-1018) */
int su2_recoupling(int kind, const int* labels, double* out);

/* Two-site SU(2) DMRG on an lx x ly cylinder (open x, periodic y) for total
 * spin 2S = totspin, keeping max_dcut multiplets (dmrg_su2::dmrg_su2,
 * dmrg_su2.cpp:31-310; flags -lx -ly -jcoup -qcoup -boundary -sec -d of
 * main_omp.cpp:42-132), with the J-Q plaquette terms when qcoup != 0
 * (dmrg_su2.cpp:200-232, 635-655).  Boundary codes (dmrg_su2.cpp:97-197): 0 square
 * lattice, periodic in x and y; 1 J1-J3; 2 J1-J2 with coup[3] on the edge
 * columns' y-bonds; 3 Shastry-Sutherland (square bonds take coup[2] = 1 - |qcoup|,
 * orthogonal dimers coup[1] = jcoup + qcoup/2).  All environments, MPS
 * projectors and Lanczos vectors stay in device memory. */
int su2_dmrg_create(su2_ctx* ctx, int lx, int ly, double jcoup, double qcoup, int bdry,
                    int totspin, int max_dcut, su2_dmrg** out);
int su2_dmrg_destroy(su2_ctx* ctx, su2_dmrg* d);
/* Warm-up growth (dmrg_su2::do_idmrg, dmrg_su2_omp.cpp:20-95); *energy is the
 * last Ritz value (the reference's gs_enr + 1000, dmrg_su2.cpp:536-538). */
int su2_dmrg_warmup(su2_ctx* ctx, su2_dmrg* d, double* energy);
/* One sweep (dmrg_su2::sweep, dmrg_su2_omp.cpp:98-157); *energy is the Ritz
 * value of its last (centre) step. */
int su2_dmrg_sweep(su2_ctx* ctx, su2_dmrg* d, double* energy);
/* Change the kept multiplet count for later steps (the reference restarts with
 * -d/-dr to ramp D, dmrg_su2_omp.cpp:974-981). */
int su2_dmrg_set_max_dcut(su2_ctx* ctx, su2_dmrg* d, int max_dcut);
/* Pin the H_eff tile variants used in every step (see su2_heff_set_variants). */
int su2_dmrg_set_heff_variants(su2_ctx* ctx, su2_dmrg* d, int stage1, int stage2);
/* Checkpoint in the reference's file format, into directory dir
 * (dmrg_su2::save_mps2 / save_ww / save_enr, dmrg_su2_omp.cpp:864-941,
 * 1070-1186): uu-{ns}-{D}-{2S}-{ky}-{exci}-{i}.dat/.bin for every site (the
 * three-leg MPS tensors with their make_cgc CG doubles), ww-...-{ns/2}.bin
 * (max_dcut*2 doubles: the last solve's singular values, zero beyond) and
 * enr-....dat (the last Ritz value minus 1000 appended, 16 digits).  Call
 * after a sweep. */
int su2_dmrg_save(su2_ctx* ctx, su2_dmrg* d, const char* dir, int ky, int exci);
/* read_mps + prepare_sweep (dmrg_su2_omp.cpp:944-1068, 160-193): replace the
 * state by the files of a run at file_dcut multiplets (0: this run's D; the
 * reference's -dr restart), rebuilding every environment; su2_dmrg_sweep then
 * continues from the centre.  *energy (nullable) receives the saved energy
 * (read_enr adds the 1000 back, 1100-1152).  Files written by the reference
 * load as well: a block whose CG is r times make_cgc is scaled by r.  On any
 * error the previous state is kept. */
int su2_dmrg_load(su2_ctx* ctx, su2_dmrg* d, const char* dir, int file_dcut, int ky, int exci,
                  double* energy);
/* Excited-state targeting (the reference's -e exci > 0: orth[j] read by
 * read_mps, dmrg_su2_omp.cpp:944-1068, its energy by read_enr, 1100-1152).
 * Adds the state saved in dir as uu/enr-{ns}-{file_dcut}-{2S}-{ky}-{exci}-*
 * (file_dcut 0: this run's max_dcut) to the states the run is kept
 * orthogonal to: the run carries the overlap environments ovlp[j]
 * (dmrg_su2_omp.cpp:327-328, 386, 470-471, 528) and every H_eff apply on
 * adjacent sites adds -E_j <phi_j|psi> phi_j (dmrg_su2.cpp:611-623).  For the
 * n-th excited state call it for exci = 0 .. n-1, before su2_dmrg_warmup (or
 * after su2_dmrg_load), and save the result with exci = n. */
int su2_dmrg_add_lower_state(su2_ctx* ctx, su2_dmrg* d, const char* dir, int file_dcut, int ky,
                             int exci);
/* level 1: one line per solve on stderr (step, sites, Ritz value, kept
 * multiplets, discarded weight, matvecs, seconds); the reference logs each
 * step to stdout (dmrg_su2_omp.cpp:113,129,145). */
int su2_dmrg_set_verbose(su2_dmrg* d, int level);
/* Ritz values of every step so far (the reference prints them per step,
 * dmrg_su2.cpp:539); *n receives the count. */
int su2_dmrg_energies(const su2_dmrg* d, double* out, int cap, int* n);
/* Per-step record, in step order: kept multiplets of the new bond (dout of
 * tensor_su2::svd, tensor_su2.cpp:1619-1752), the summed squared singular
 * values its global cut drops (tensor_su2.cpp:1686-1694; the reference keeps
 * only the kept values, wout) and the H_eff applies of the Lanczos run
 * (lanczos_su2.cpp:237-338).  Any output may be NULL; *n receives the count. */
int su2_dmrg_history(const su2_dmrg* d, int* kept, double* discarded, int* matvecs, int cap,
                     int* n);
/* Matvec arithmetic of later solves: bits = 64 (default, the reference's
 * FP64 throughout) or 32 (mixed precision): the Lanczos Krylov matvecs run on
 * FP32 MFMA with FP32 accumulation (su2_heff_set_precision), while the
 * Ritz-vector check of lanczos_su2::check_eigenvector (lanczos_su2.cpp:
 * 320-332) runs in FP64 and its Rayleigh quotient <psi|H psi> is the step's
 * energy (an FP64 correction of the Ritz value, exact to second order in the
 * Ritz vector's error).  Lanczos vectors and coefficients, the tridiagonal
 * eigensolve, the SVD/truncation and the environment updates stay FP64.
 * Solves whose matvec has fewer than min_flops algorithmic FLOPs stay FP64
 * (min_flops < 0: the default 2e8, below which the FP32 copies cost more
 * than they save).  A mixed run schedules 32 for the early sweeps and 64 for
 * the final ones. */
int su2_dmrg_set_matvec_precision(su2_ctx* ctx, su2_dmrg* d, int bits, double min_flops);
/* The two-site H_eff of the centre bond (il = ns/2 - 1, ir = ns/2) of the
 * current state, built from the run's environments as every solve builds it
 * (dmrg_su2.cpp:543-664, without excited-state penalties), and the state's
 * wavefunction in that basis (A_j V_j^T per sector, the factors the last
 * solve left, dmrg_su2_omp.cpp:135-150).  Call after su2_dmrg_warmup or
 * su2_dmrg_sweep; destroy both with su2_heff_destroy / su2_tensor_destroy.
 * For measurements on the real state (energy, variance, hermiticity) and
 * timing the matvec on the run's own sector dimensions. */
int su2_dmrg_centre(su2_ctx* ctx, su2_dmrg* d, su2_heff** plan, su2_tensor** psi);
/* H_eff launches of each solve replayed from a HIP graph (on = 1, default)
 * or launched directly (0); the results are the same. */
int su2_dmrg_set_graph(su2_ctx* ctx, su2_dmrg* d, int on);

/* Several DMRG runs of one context advanced in lockstep (no counterpart in
 * the reference, which runs one job per process).  Each tick takes the next
 * two-site solve of every unfinished run (dmrg_su2_omp.cpp:20-157), runs all
 * their Lanczos solves (lanczos_su2.cpp:125-195) together -- one grouped
 * launch per GEMM stage for all of their H_eff applies each round, captured
 * in a HIP graph and replayed, one host wait per round -- then one batched
 * SVD over all their wavefunctions, then each run's environment update.
 * Every problem keeps its own K split and every solve its own reduction
 * order: each run's energies are bit for bit those it has alone. */
int su2_batch_create(su2_ctx* ctx, su2_batch** out);
int su2_batch_destroy(su2_ctx* ctx, su2_batch* b);
/* Adds run d (same context, not owned; keep it until the batch is done with
 * it): a warm-up, then nsweep sweeps with max_dcut[k] multiplets and matvec
 * precision bits[k] (64 or 32 with min_fp32_flops as in
 * su2_dmrg_set_matvec_precision; bits NULL: 64); the warm-up uses entry 0.
 * stop > 0 ends the run once two successive sweeps differ by less than stop
 * (main_omp.cpp:197).  *id: the run's index. */
int su2_batch_add(su2_ctx* ctx, su2_batch* b, su2_dmrg* d, int nsweep, const int* max_dcut,
                  const int* bits, double min_fp32_flops, double stop, int* id);
/* One lockstep step; *stepped = number of runs that took a step (0: all done). */
int su2_batch_tick(su2_ctx* ctx, su2_batch* b, int* stepped);
/* Ticks until every run has finished. */
int su2_batch_run(su2_ctx* ctx, su2_batch* b);
/* Runs n batches to completion at once, one host thread each ("lanes"); batch
 * i must belong to ctxs[i] and every context must be distinct (one stream per
 * lane).  One lane's host work (plan builds, environment scatters, its waits
 * for Lanczos scalars and Jacobi sweeps) then overlaps the others' device
 * work.  Each run's results are bit for bit those it gets in one lane.  On
 * failure the other lanes stop at their next tick; the failing lane's message
 * is in su2_last_error(ctxs[i]) and su2_last_error(NULL).  No counterpart in
 * the reference (one job per process). */
int su2_batch_run_lanes(su2_ctx* const* ctxs, su2_batch* const* bs, int n);
int su2_batch_set_graph(su2_ctx* ctx, su2_batch* b, int on);
/* Energy after each finished phase of run id (warm-up first); *finished. */
int su2_batch_energies(const su2_batch* b, int id, int* finished, double* out, int cap, int* n);
/* out[0..7] = ticks, HIP graphs built, graph replays, s preparing solves,
 * s in Lanczos, device s of the batched applies, s in SVD, s completing. */
int su2_batch_stats(const su2_batch* b, double* out, int cap);

/* Request loop with continuous batching on one device (main_omp.cpp:42-202
 * per job, dmrg_su2_omp.cpp:944-1068 for excited levels).  Jobs are admitted
 * in arrival order once they have arrived, one of max_slots slots is free
 * and their estimated device footprint (su2_job_footprint) fits what the
 * admitted jobs leave of budget_bytes.  Admitted jobs keep their MPS and
 * environments on the device and advance together, one solve each per
 * su2_batch tick; a finished job frees its slot and memory at once and a job
 * arriving meanwhile joins at the next tick.  An excited job of level k runs
 * levels 0..k in its slot, saving each in the reference's file format under
 * workdir/job{id} for the next (su2_dmrg_add_lower_state).  Each job's
 * result is bit for bit what it gives served alone. */
int su2_server_create(su2_ctx* ctx, double budget_bytes, int max_slots, const char* workdir,
                      su2_server** out);
int su2_server_destroy(su2_ctx* ctx, su2_server* s);
/* arrival_s: seconds after the start of su2_server_run. */
int su2_server_submit(su2_ctx* ctx, su2_server* s, const su2_job* job, double arrival_s,
                      int* id);
/* Lanes (1..8, default 1): the slots are split over this many lanes, each on
 * its own context (stream) of ctx's device with its own batch driver and host
 * thread, taking arrived jobs in arrival order under the one memory budget.
 * A job's energy does not depend on the lane count.  Set before su2_server_run. */
int su2_server_set_lanes(su2_ctx* ctx, su2_server* s, int lanes);
/* Serves every submitted job to completion (an idle device waits for the
 * next arrival); verbose 1 logs admissions and completions on stderr. */
int su2_server_run(su2_ctx* ctx, su2_server* s, int verbose);
/* Final energy (Ritz value of the target level), times[3] = arrival, admission
 * and completion in seconds from the start of the run, estimated footprint. */
int su2_server_result(const su2_server* s, int id, double* energy, double* times,
                      double* footprint);
/* out[0..2] = wall seconds of the run, ticks, most jobs admitted at once. */
int su2_server_stats(const su2_server* s, double* out, int cap);
/* Estimated device bytes of a job (environments of every site, a step's
 * H_eff bundles and Lanczos vectors, SVD work space; more for excited
 * levels, which hold their lower states). */
int su2_job_footprint(const su2_job* job, double* bytes);

/* Timing counters: out[0..16] = environment s, Lanczos s, SVD s, projection
 * s, start-vector s, device matvec s, algorithmic H_eff FLOPs, matvecs, steps,
 * last centre bond dimension, s inside the one-sided Jacobi SVD, its sweeps
 * summed over solves, FP32 matvecs, their device s, their algorithmic FLOPs,
 * the largest |Ritz value - FP64 Rayleigh quotient| of a mixed solve, SVDs
 * accepted at the Jacobi sweep ceiling (60 sweeps; columns orthogonal to
 * 1e-6 rather than 1e-13, singular values then good to ~1e-12 relative;
 * beyond 1e-6 the step fails with SU2_ERR_NUMERIC).  cap bounds what is
 * written. */
int su2_dmrg_stats(const su2_dmrg* d, double* out, int cap);

#ifdef __cplusplus
}
#endif

#endif /* SU2DMRG_H */
