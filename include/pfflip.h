/*
 * pfflip.h -- C ABI of the MI355X (gfx950) PF-FLIP time-step library
 * (libpfflip_hip.so).
 *
 * The reference (a Python + Warp package) has no FFI layer: its boundary is
 * the Python object API of `Simulator` / `MGPCG`, whose stages launch Warp
 * kernels.  Each entry point below replaces the stage (or kernel group) cited
 * next to it, as `path:line` in the reference repository.  The Python
 * drop-in (`Simulator`, `MGPCG`, `apply_A`) binds these with ctypes
 * (implementation-of-adaptive-phase-field-flip_amd/_lib.py).
 *
 * Conventions
 *  - Return 0 on success, a negative PF_ERR_* for invalid arguments, or a
 *    positive hipError_t.  pf_last_error() gives a thread-local message.
 *  - All arrays are caller-owned device memory, C-contiguous with the last
 *    index fastest.  Cell (i,j,k) is element (i*ny + j)*nz + k.  Face arrays
 *    have shapes (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1).  Particle
 *    positions/velocities are float32 (n,3) arrays of structures.
 *  - Every call is asynchronous on `stream` (a hipStream_t), except
 *    pf_mgpcg_solve and pf_read_scalar*, which block until their result is
 *    known.  No entry point allocates device memory unless it says so.
 *  - No exceptions cross the ABI.
 */
#ifndef PFFLIP_H
#define PFFLIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_OK 0
#define PF_ERR_INVALID (-1)   /* bad pointer, size or flag combination      */
#define PF_ERR_SIZE (-2)      /* a size exceeds what a kernel supports      */
#define PF_ERR_UNSUPPORTED (-3)

#define PF_MAX_LEVELS 16
#define PF_TILE 8 /* P2G tile edge in cells; binning keys are tile-major */

typedef void* pf_stream_t; /* hipStream_t */

/* ---------------------------------------------------------------- common */
const char* pf_last_error(void);
int pf_abi_version(void);

/* Fill `n` 32-bit words with `value` (a kernel, safe under graph capture). */
int pf_fill_u32(void* dst, int64_t n, uint32_t value, pf_stream_t stream);

/* Read one device float / double / int into host memory (blocks). */
int pf_read_scalars(const void* dev_src, void* host_dst, size_t bytes, pf_stream_t stream);

/* ------------------------------------------------ particle binning + sort
 * Build-added: the reference splats with per-particle global atomics in
 * particle order (pfflip/kernels.py:100-209) and has no sort.  Keys feed
 * pf_p2g.  For every particle with pstate == 0:
 *   cell = clamp(floor(x / dx), 0, n-1) per axis (FP32 division),
 *   key  = tile_id * PF_TILE^3 + local_id,
 *   tile_id = (tx*nty + ty)*ntz + tz, local_id = (li*T + lj)*T + lk.
 * Particles are visited in `order` (a permutation of 0..n-1, normally the
 * previous step's `perm`; NULL = index order); keys[i] and rank[i] belong
 * to the i-th visited particle (key 0xFFFFFFFF = not a FLIP particle).
 * On return `perm` is a permutation of all n particles: the FLIP particles
 * sorted by key, then the others.  With `stable` != 0 ties are broken by
 * particle index (a unique order); otherwise the order within a key depends
 * on atomic timing (pf_p2g's sums do not depend on it).  cell_start (nkeys + 2 entries, nkeys = pf_bin_nkeys()) gives
 * key k the range cell_start[k] .. cell_start[k+1]; cell_start[nkeys] is
 * the number of FLIP particles and cell_start[nkeys+1] = n.
 * Buffers: keys, rank, perm: n entries; scan_tmp: pf_bin_scan_tmp_bytes().
 * *n_active (device int32, may be NULL) also receives the FLIP count. */
int64_t pf_bin_nkeys(int nx, int ny, int nz);
size_t pf_bin_scan_tmp_bytes(int nx, int ny, int nz);
int pf_bin_sort(const float* px, const int32_t* pstate, int64_t n, int nx, int ny, int nz,
                float dx, const uint32_t* order, uint32_t* keys, uint32_t* rank,
                uint32_t* perm, int32_t* cell_start, void* scan_tmp, int32_t* n_active,
                int32_t stable, pf_stream_t stream);

/* pf_bin_sort without a scan_tmp argument and without any allocation (no
 * reference counterpart: the reference has no sort).  For callers that
 * cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  The scan buffer is a
 * static device array of the library, one per GPU, held until the process
 * ends, sized for at most PF_BIN_NOALLOC_MAX_KEYS binning keys
 * (pf_bin_nkeys(); 2^21 keys = a 128^3 grid, about 8.4 MB); a larger grid
 * returns PF_ERR_SIZE.  Results equal pf_bin_sort's: keys, cell_start and
 * *n_active bit for bit, and with stable != 0 the FLIP particles' part of
 * perm too (in both, the particles outside the bins close perm in
 * atomic-timing order).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (two streams, two threads) race and must be serialised by the
 * caller.  The array's counts are cleared by a fill kernel at the start of
 * each call. */
#define PF_BIN_NOALLOC_MAX_KEYS 2097152
int pf_bin_sort_noalloc(const float* px, const int32_t* pstate, int64_t n, int nx, int ny, int nz,
                        float dx, const uint32_t* order, uint32_t* keys, uint32_t* rank,
                        uint32_t* perm, int32_t* cell_start, int32_t* n_active,
                        int32_t stable, pf_stream_t stream);

/* ---------------------------------------------------------------- P2G
 * Replaces Simulator._p2g's zeroing + p2g_face x3 + p2g_cell
 * (pfflip/sim.py:205-228, pfflip/kernels.py:100-209): eq. (6) weights
 * w = max(1 - d^2/r^2, 0)^3 with r = rp (= dx), mass += w*m,
 * mom_a += w*m*v_a, cell_mass += w*m, cell_wsum += w (if non-NULL).
 * Sums are formed in 64-bit fixed point (one power-of-two scale per
 * quantity, sized from `bounds` so they cannot overflow), so the result does
 * not depend on summation order; each output is that sum rounded to FP32.
 * bounds: device floats [max|v|, max m, max m*max_a|v_a|] over the FLIP
 * particles, as written by pf_max_speed(..., pmass, ...).  Every output
 * element is written (no pre-zeroing needed).  Consumes perm/cell_start from
 * pf_bin_sort.  `stage` is scratch of pf_p2g_stage_bytes(nx, ny, nz,
 * cell_wsum != NULL) bytes. */
size_t pf_p2g_stage_bytes(int nx, int ny, int nz, int with_wsum);
int pf_p2g(const float* px, const float* pv, const float* pmass,
           const uint32_t* perm, const int32_t* cell_start,
           int nx, int ny, int nz, float dx, float rp, const float* bounds,
           float* mass_x, float* mass_y, float* mass_z,
           float* mom_x, float* mom_y, float* mom_z,
           float* cell_mass, float* cell_wsum, void* stage, size_t stage_bytes,
           pf_stream_t stream);

/* ------------------------------------------ phase field / beta / forces
 * One fused grid pass; `flags` selects the reference stages it performs,
 * applied per face in the reference's order:
 *   PF_PH_VEL_FROM_MOM  vel = mom/mass if mass > mass_eps else 0, valid
 *                       (face_velocity_from_momentum, kernels.py:212-228);
 *                       without it `vel` is read.
 *   PF_PH_PHASE         phi_cell, is_liquid (cell_phase, kernels.py:242-255),
 *                       phi_face, beta (face_beta, kernels.py:258-311)
 *   PF_PH_ZERO          walls and solid faces: beta = 0, vel = 0
 *                       (kernels.py:314-326, 601-627)
 *   PF_PH_STAR          vel_star = vel after zeroing (sim.py:176-177)
 *   PF_PH_GRAVITY       vel += g_a*dt, then re-zero (sim.py:270-285)
 *   PF_PH_RHS           rhs from the final vel (build_rhs, kernels.py:338-363)
 *   PF_PH_TAGS_IN       with PF_PH_PHASE: is_liquid is an input (liquid tags
 *                       replayed from another run, e.g. to difference a
 *                       simulation along the branch its adjoint follows);
 *                       phi_cell is still written
 * `solid` may be NULL (no obstacles).  PF_PH_RHS together with PF_PH_ZERO or
 * PF_PH_GRAVITY needs PF_PH_VEL_FROM_MOM: a cell's RHS reads its high faces,
 * which the same pass rewrites when vel is the input (PF_ERR_INVALID
 * otherwise). */
#define PF_PH_VEL_FROM_MOM 1
#define PF_PH_PHASE 2
#define PF_PH_ZERO 4
#define PF_PH_STAR 8
#define PF_PH_GRAVITY 16
#define PF_PH_RHS 32
#define PF_PH_TAGS_IN 64

typedef struct pf_phase_args {
    int32_t nx, ny, nz, flags;
    const float* mass[3];
    const float* mom[3];
    float* vel[3];
    float* vel_star[3];
    int32_t* valid[3];      /* may be NULL */
    const float* cell_mass;
    const float* cell_wsum; /* read only when k_div > 0 */
    float* phi_cell;
    int32_t* is_liquid;
    float* phi_face[3];
    float* beta[3];
    const int32_t* solid; /* may be NULL */
    float* rhs;
    float mass_eps, rho_min, alpha_phi, rho_0l, rho_l, rho_g;
    float gravity[3];
    float dt, dx, rho_tilde_0, k_div;
} pf_phase_args;

int pf_phase_beta_forces(const pf_phase_args* args, pf_stream_t stream);
/* build_rhs (kernels.py:338-363) alone: rhs = -(dx/dt) sum_a (u_a,hi -
 * u_a,lo), minus dx^2 k_div max((cell_wsum - rho_tilde_0)/rho_tilde_0, 0)/dt
 * when k_div > 0 (cell_wsum may be NULL otherwise).  The same pass as
 * pf_phase_beta_forces with flags = PF_PH_RHS; vel is only read. */
int pf_build_rhs(int32_t nx, int32_t ny, int32_t nz, const float* const vel[3], const float* cell_wsum,
                 float* rhs, float dx, float dt, float k_div, float rho_tilde_0, pf_stream_t stream);

/* ---------------------------------------------------------------- MG-PCG
 * Replaces MGPCG (pfflip/poisson.py:226-409).  Levels halve while
 * min(n)//2 >= min_dim and all dims are even (poisson.py:256-263).
 * precision: 0 = FP32 V-cycle (the reference's arithmetic), 1 = BF16 face
 * coefficients and smoother iterates with FP32 accumulation, 2 = FP16.  CG
 * vectors are FP32; CG dot products, alpha/beta, the stopping test and the
 * null-space (mean) projections accumulate in FP64 in every mode. */
typedef struct pf_mg_level {
    int32_t nx, ny, nz, _pad;
    float* bx;
    float* by;
    float* bz;
    float* diag;
    float* p;
    float* b;
    /* second iterate buffer of the fused smoother (which reads one buffer and
     * writes the other) on the levels below small_from; in 16-bit mode it
     * holds 16-bit values.  Level 0's `p` is the FP32 mode's second buffer. */
    float* r;
    /* reduced-precision copies, used when precision != 0 on the levels below
     * small_from; level 0's p_h (the preconditioned vector) always */
    uint16_t* bx_h;
    uint16_t* by_h;
    uint16_t* bz_h;
    uint16_t* p_h;
} pf_mg_level;

typedef struct pf_mg {
    int32_t nlev, pre, post, coarsest;
    float omega;
    int32_t precision;
    int32_t small_from; /* out: first level solved inside one workgroup (set by pf_mg_setup) */
    int32_t options;    /* PF_MG_* bits, 0 = defaults */
    pf_mg_level lev[PF_MAX_LEVELS];
    float* rhs; /* CG work vectors, level-0 sized (rhs and q hold the
                   alternating search directions) */
    float* res;
    float* z;
    float* q;
    float* Ap;
    void* scalars; /* pf_mg_scalar_bytes() of device memory */
    /* solver host state from pf_mgpcg_state_create (the pinned flag the
     * residual tests publish to and the HIP graphs captured for these
     * buffers), owned by the caller;
     * NULL: each pf_mgpcg_solve creates and frees its own, without graphs */
    void* host_state;
} pf_mg;

/* Smooth with two half-sweep launches per sweep instead of the fused
 * full-sweep kernel (same results bit for bit; for testing and timing). */
#define PF_MG_UNFUSED_SMOOTHER 1
/* Form the flexible CG's inner products after the preconditioner in a
 * kernel of their own; by default the V-cycle's last level-0 sweep forms
 * them from the values it holds (the same sums in another order: results
 * agree to rounding; for testing and timing). */
#define PF_MG_SEPARATE_DOTS 2
/* Run the small levels below level 0 (up to 32^3 cells) with one kernel per
 * sweep like the larger ones; by default each of their smoothing phases
 * (pre-sweeps with the restriction, prolongation with the post-sweeps) is
 * one kernel over LDS-resident tiles (same results bit for bit; for testing
 * and timing). */
#define PF_MG_SPLIT_LEVELS 4

size_t pf_mg_scalar_bytes(void);
/* Validate the level table and choose small_from. */
int pf_mg_setup(pf_mg* mg);
/* build_hierarchy (poisson.py:295-309): copy beta into level 0 unless the
 * pointers already alias it, diag per level, Galerkin coarse faces = sum of
 * the 4 fine faces (poisson.py:130-163), reduced-precision copies. */
int pf_mg_build(const pf_mg* mg, const float* bx, const float* by, const float* bz,
                pf_stream_t stream);
/* precondition (poisson.py:345-352): z = V-cycle(r) minus its mean. */
int pf_mg_precondition(const pf_mg* mg, const float* r, float* z, pf_stream_t stream);
/* _vcycle (poisson.py:323-343): z = one V-cycle from a zero guess on the
 * right-hand side r (the coarsest level's mean removal included, no final
 * mean removal).  FP32 hierarchies only; r, z and lev[0].p distinct. */
int pf_mg_vcycle(const pf_mg* mg, const float* r, float* z, pf_stream_t stream);
/* Host state of one solver for pf_mg.host_state: created on the current
 * device, destroyed with pf_mgpcg_state_destroy (no reference counterpart;
 * MGPCG.__init__'s buffers, poisson.py:249-275, hold no host state).  One
 * state serves one thread at a time. */
int pf_mgpcg_state_create(void** out);
int pf_mgpcg_state_destroy(void* state);
/* solve (poisson.py:355-409): flexible PCG, Polak-Ribiere beta clamped at 0,
 * stop when max|r|/max|b| < tol (recursive residual, b mean-removed).
 * Writes x (mean removed); *it_out / *rel_out are host pointers.  Blocks:
 * every residual test publishes (tests so far, done) to a pinned host int,
 * which the host polls, while the GPU still has queued work, to stop
 * enqueueing iterations, and the result is published the same way (no
 * events, no copies, no stream synchronisation).  The library keeps no state of its own between
 * calls: graphs and flags live in mg->host_state. */
int pf_mgpcg_solve(const pf_mg* mg, const float* b, float* x, double tol, int32_t max_iter,
                   int32_t* it_out, double* rel_out, pf_stream_t stream);
/* Measurement hook (no reference counterpart): average milliseconds of one
 * smoother sweep of HBM level `level` over `reps` sweeps, timed with HIP
 * events on `stream`: one fused full-sweep kernel, or with
 * PF_MG_UNFUSED_SMOOTHER in mg->options one red-black half-sweep kernel.
 * Overwrites that level's iterate buffers (use after a solve).  Blocks. */
int pf_mg_time_smoother(const pf_mg* mg, int32_t level, int32_t reps, float* ms_out,
                        pf_stream_t stream);
/* apply_A (poisson.py:54-88): out = diag*p - sum_f beta_f p_nbr with
 * neighbour indices clamped at the walls.  diag may be NULL (then it is
 * summed from the six faces in compute_diag's order, poisson.py:40-51). */
int pf_apply_A(int nx, int ny, int nz, const float* p, const float* bx, const float* by,
               const float* bz, const float* diag, float* out, pf_stream_t stream);

/* ---------------------------------------------------------------- projection
 * Replaces _project + _max_divergence (pfflip/sim.py:301-314,
 * kernels.py:366-400): on interior faces with beta != 0,
 * vel -= dt*beta*(p_i - p_{i-1})/dx.  If dvel[a] and vel_star[a] are
 * non-NULL, dvel = vel - vel_star (field_difference, kernels.py:499-502).
 * If max_div (device float) is non-NULL it receives max |div(vel)/dx| of the
 * projected field.  Pass dt = 0 to only measure the divergence. */
int pf_project(const float* p, float* const beta[3], float* const vel[3],
               float* const vel_star[3], float* const dvel[3],
               int nx, int ny, int nz, float dx, float dt, float* max_div,
               pf_stream_t stream);

/* ---------------------------------------------------------------- extrapolation
 * Replaces _extrapolate_liquid_velocity (pfflip/sim.py:329-341,
 * kernels.py:406-462): seed faces touching a liquid cell, then `sweeps`
 * Jacobi sweeps of the mean of known face-index neighbours.  Result in
 * vel_liq; tmp (3 face arrays) and known/known2 (3+3 uint8 face arrays) are
 * scratch. */
int pf_extrapolate(float* const vel[3], const int32_t* is_liquid, float* const vel_liq[3],
                   float* const tmp[3], uint8_t* const known[3], uint8_t* const known2[3],
                   int nx, int ny, int nz, int sweeps, pf_stream_t stream);

/* ---------------------------------------------------------------- G2P / advection
 * Replaces _g2p (sim.py:316-327, kernels.py:468-502) and _advect
 * (sim.py:408-424, kernels.py:508-549, 567-595, 630-679).  flags:
 *   PF_GA_G2P     v = a(v + I(dvel)) + (1-a) I(vel), a per phase (eq. 12)
 *   PF_GA_ADVECT  Ralston RK3 through vel_liq (liquid) / vel (air), clamp
 *   PF_GA_REFLECT clamp into [lo,hi] and kill the inward component
 *   PF_GA_SOLIDS  push particles out of solid cells (needs `solid`)
 * Per-particle phase comes from ptype/pstate, or from a packed byte per
 * particle in pflag when non-NULL (bit 0 liquid, bit 1 pstate == 0,
 * bit 2 pstate == 2). */
#define PF_GA_G2P 1
#define PF_GA_ADVECT 2
#define PF_GA_REFLECT 4
#define PF_GA_SOLIDS 8

typedef struct pf_g2p_args {
    float* px;
    float* pv;
    const int32_t* ptype;
    const int32_t* pstate;
    const uint8_t* pflag;
    int64_t n;
    int32_t nx, ny, nz, flags;
    const float* vel[3];
    const float* dvel[3];
    const float* vel_liq[3];
    const int32_t* solid;
    float dx, dt, alpha_liquid, alpha_air;
    float lo[3], hi[3];
} pf_g2p_args;

int pf_g2p_advect(const pf_g2p_args* args, pf_stream_t stream);

/* The same stages for particles binned by pf_bin_sort at their current
 * positions (perm, cell_start, n_active from that call; pstate unchanged
 * since): one workgroup per 8^3 tile samples LDS copies of the tile's face
 * fields, then particles outside the bins (pstate 1) get the reflection /
 * solids part.  Results equal pf_g2p_advect's.  If bounds_out is non-NULL
 * (then pmass is required) it receives what pf_max_speed(pv, pmass, pstate)
 * returns after the call, computed from the updated particles in the same
 * pass.  Replaces the same reference kernels (kernels.py:468-549,552-564,
 * 567-679). */
int pf_g2p_advect_binned(const pf_g2p_args* args, const int32_t* perm, const int32_t* cell_start,
                         const int32_t* n_active, const float* pmass, float* bounds_out,
                         pf_stream_t stream);

/* particle_max_speed (kernels.py:552-564): out[0] = max |v| over
 * pstate == 0 (device floats, overwritten).  With pmass non-NULL also
 * out[1] = max m and out[2] = max m*max_a|v_a| (the P2G bounds). */
int pf_max_speed(const float* pv, const float* pmass, const int32_t* pstate, int64_t n,
                 float* out, pf_stream_t stream);

/* ---------------------------------------------------------------- droplets
 * Replaces _droplets (pfflip/sim.py:343-390, droplets.py:67-278).  Per
 * particle, in this order: classify_escaped (liquid FLIP particle where the
 * interpolated cell phi < 1/2 - delta_phi -> droplet with a log-normal radius;
 * air FLIP particle where phi > 1/2 + delta_phi -> dead; droplet where
 * phi > 1/2 + delta_phi -> FLIP), emit_spray_in_shear (only if shear is
 * non-NULL), advance_droplets (gravity + sphere drag against vel, per-droplet
 * sub-steps, optional evaporation, death on leaving [lo, hi]).  seed is the
 * reference's cfg.seed + step_index.  counters: device int32[5], added to
 * (not cleared): emitted, bubbles, merged, spray, droplets alive after the
 * stage. */
typedef struct pf_droplet_args {
    float* px;
    float* pv;
    const int32_t* ptype;
    int32_t* pstate;
    float* pdrop_r;
    int64_t n;
    int32_t nx, ny, nz;
    int32_t seed;
    const float* phi;
    const float* shear;
    const float* vel[3];
    float dx, dt, delta_phi;
    float r_mu, r_sigma, r_min, r_max;
    float spray_threshold, spray_rate;
    float gravity[3];
    float rho_g, rho_l, cd, evap_k;
    float lo[3], hi[3];
    int32_t max_substeps;
    int32_t* counters;
} pf_droplet_args;

int pf_droplets(const pf_droplet_args* args, pf_stream_t stream);

/* velocity_gradient_magnitude (droplets.py:156-196): out (cells) = |grad u|
 * from the MAC faces with the reference's seven terms. */
int pf_velocity_gradient(const float* const vel[3], float* out, int nx, int ny, int nz, float dx,
                         pf_stream_t stream);

/* rasterize_spray (droplets.py:308-351, sim.py:392-401): adds droplet volume
 * trilinearly into density (FP16 bits, sx*sy*sz cells of size dx_spray,
 * 4-byte aligned), each contribution rounded to FP16 and added atomically.
 * The atomics update whole 4-byte words, so when sx*sy*sz is odd the
 * allocation must hold one more FP16 after the last cell (it is added 0). */
int pf_rasterize_spray(const float* px, const int32_t* pstate, const float* pdrop_r, int64_t n,
                       uint16_t* density, int sx, int sy, int sz, float dx_spray, pf_stream_t stream);

/* ---------------------------------------------------------------- MSBG
 * Multiresolution sparse block grid (pfflip/msbg.py): nb blocks of one
 * geometric size; a level-l block holds (n0 >> l)^3 cells of spacing
 * h0 * 2^l.  The device arrays are the reference's layout (msbg.py:47-65):
 * level[nblk] (-1 inactive), slot[nblk] (-1 inactive), blk[nslots],
 * blev[nslots], coff[nslots] (first cell address), cslot[ncells]; blocks in
 * slots are Morton ordered; cell (i,j,k) of a slot at coff + (i n + j) n + k.
 * Face coefficients: bface[3 * addr + axis] is the cell's low face. */
typedef struct pf_msbg_grid {
    int32_t nbx, nby, nbz, n0, sh0, nslots, ncells;
    float h0;
    const int32_t* level;
    const int32_t* slot;
    const int32_t* blk;
    const int32_t* blev;
    const int32_t* coff;
    const int32_t* cslot;
} pf_msbg_grid;

/* apply_A / compute_diag (mr_operator.py:74-180): out = A p, or with
 * want_diag the diagonal (p unused). */
int pf_msbg_apply(const pf_msbg_grid* g, const float* p, const float* bface, float* out, int32_t want_diag,
                  pf_stream_t stream);
/* residual (mr_operator.py:183-190): r = b - A p. */
int pf_msbg_residual(const pf_msbg_grid* g, const float* p, const float* b, const float* bface, float* r,
                     pf_stream_t stream);
/* rbgs_sweep (mr_operator.py:193-212): one colour of weighted Gauss-Seidel,
 * colour from the finest-coordinate checkerboard, cells of that colour
 * updated from pin (all of them read the values before the half-sweep),
 * the others copied; pout != pin. */
int pf_msbg_rbgs(const pf_msbg_grid* g, const float* pin, float* pout, const float* b, const float* bface,
                 const float* diag, int32_t colour, float omega, pf_stream_t stream);
/* Adaptive relaxation (adaptive_relax.py:84-160).  pf_msbg_rbgs_active:
 * rbgs_sweep_active, one colour over the blocks with active[slot] != 0 from
 * pin into pout (all cells written), delta[c] = |update| of the updated
 * cells (the caller zeroes delta before each sweep).  pf_msbg_relax_update:
 * block_effective_residual (res[slot] = max delta diag, 0 if inactive);
 * with set_theta, theta[0] = exp(mean + theta_sigma std) of log res over
 * the positive entries (+inf if none); update_active (active_next = blocks
 * with res > theta and their face neighbours). */
int pf_msbg_rbgs_active(const pf_msbg_grid* g, const float* pin, float* pout, const float* b, const float* bface,
                        const float* diag, const int32_t* active, int32_t colour, float omega, float* delta,
                        pf_stream_t stream);
int pf_msbg_relax_update(const pf_msbg_grid* g, const float* delta, const float* diag, const int32_t* active,
                         float* res, float* theta, int32_t set_theta, float theta_sigma, int32_t* active_next,
                         pf_stream_t stream);
/* The coarsest level's smoothing of _vcycle (mr_multigrid.py:216-219): all
 * `sweeps` red-black sweeps in one workgroup from LDS, equal bit for bit to
 * as many pf_msbg_rbgs pairs; p is read and overwritten.  At most
 * pf_msbg_coarsest_max_cells() cells (the current device's LDS per
 * workgroup). */
int32_t pf_msbg_coarsest_max_cells(void);
int pf_msbg_coarsest_smooth(const pf_msbg_grid* g, float* p, const float* b, const float* bface, const float* diag,
                            int32_t sweeps, float omega, pf_stream_t stream);
/* The coarsest level's smoothing as a matrix (no counterpart: the reference
 * sweeps): `sweeps` red-black sweeps from zero plus mean removal are linear
 * in b, so pf_msbg_coarse_op writes op[j * n + t] = entry (t, j) of that
 * map (FP64, one workgroup per column, n = g->ncells <=
 * pf_msbg_coarse_op_max_cells()), once per set of face coefficients; then
 * pf_msbg_coarse_apply gives p = op b (FP64 sums rounded to float) per
 * V-cycle. */
int32_t pf_msbg_coarse_op_max_cells(void);
int pf_msbg_coarse_op(const pf_msbg_grid* g, const float* bface, const float* diag, int32_t sweeps, float omega,
                      double* op, pf_stream_t stream);
int pf_msbg_coarse_apply(const double* op, const float* b, float* p, int32_t n, pf_stream_t stream);
/* coarsen_beta (mr_multigrid.py:52-89): coarse face = mean of the 2x2 fine
 * faces it covers. */
int pf_msbg_coarsen_beta(const pf_msbg_grid* coarse, const pf_msbg_grid* fine, const float* bfine, float* bcoarse,
                         pf_stream_t stream);
/* restrict_sum (mr_multigrid.py:92-99): bcoarse (overwritten) = sum of the
 * fine values each coarse cell covers. */
int pf_msbg_restrict(const pf_msbg_grid* fine, const pf_msbg_grid* coarse, const float* rfine, float* bcoarse,
                     pf_stream_t stream);
/* prolong_add (mr_multigrid.py:102-109): pfine += pcoarse of the covering
 * coarse cell. */
int pf_msbg_prolong_add(const pf_msbg_grid* fine, const pf_msbg_grid* coarse, const float* pcoarse, float* pfine,
                        pf_stream_t stream);
/* cell_center / cell_volume (msbg.py:144-153, mr_operator.py:215-220):
 * centre (ncells x 3), volume, interior (1 if all six neighbours exist);
 * each may be NULL. */
int pf_msbg_geometry(const pf_msbg_grid* g, float* centre, float* volume, int32_t* interior, pf_stream_t stream);
/* hits[c] (overwritten) = number of finest coordinates locate() resolves to
 * cell c (the addressing check of tests/test_msbg.py:69-89). */
int pf_msbg_cover(const pf_msbg_grid* g, int32_t* hits, pf_stream_t stream);

/* Vector kernels of the MSBG solver (mr_solver.py:27-69).  pf_vec_reduce:
 * work[0] = sum x (op 0), <x, y> (op 1) or max |x| (op 2) in FP64, combined
 * in a fixed order; work holds pf_vec_reduce_bytes() bytes.  pf_vec_axpby:
 * y = a x + b y, or y = b y + a when x is NULL. */
size_t pf_vec_reduce_bytes(void);
int pf_vec_reduce(int32_t op, const float* x, const float* y, int64_t n, double* work, pf_stream_t stream);
int pf_vec_axpby(float a, const float* x, float b, float* y, int64_t n, pf_stream_t stream);
/* Device-resident scalars of the flexible PCG (mr_solver.py:118-165), so a
 * solve reads one state record per iteration instead of every reduction.
 * pf_vec_axpby_dev: y = a x + b y with a = a_sign * st[ia] (a = a_sign
 * when ia < 0) and b = st[ib] (1 when ib < 0), rounded to float; a no-op
 * once flag[0] != 0 (flag may be NULL).  pf_vec_remove_mean: x -= mean(x)
 * as _remove_mean (mr_multigrid.py:181-184), the FP64 sum kept on the
 * device (work as for pf_vec_reduce).  pf_mrcg_scalar: st = {rz, alpha,
 * beta, rel, bnorm, tol, qAq, flag[0] as a double}; op 0 (init: bnorm = *a, tol), 1 (rz = *a),
 * 2 (alpha = rz / *a, or flag 1 when *a <= 0), 3 (rel = *a / bnorm, flag 2
 * below tol), 4 (beta = max((*a - *b) / rz, 0), rz = *a, or flag 3 when
 * rz == 0 or *a <= 0); ops 2-4 do nothing once flag[0] != 0. */
int pf_vec_axpby_dev(const double* st, int32_t ia, float a_sign, const float* x, int32_t ib, float* y, int64_t n,
                     const int32_t* flag, pf_stream_t stream);
int pf_vec_remove_mean(float* x, int64_t n, double* work, pf_stream_t stream);
int pf_mrcg_scalar(int32_t op, double* st, int32_t* flag, const double* a, const double* b, double tol,
                   pf_stream_t stream);
/* _jacobi (mr_solver.py:64-69): z = r / diag where diag > 0, else 0. */
int pf_vec_jacobi(const float* r, const float* diag, float* z, int64_t n, pf_stream_t stream);

/* ---------------------------------------------------------------- adaptive step
 * The adaptive PF-FLIP step on an MSBG (pfflip/mr_sim.py:203-310 with the
 * kernels of pfflip/mr_kernels.py and pfflip/particles.py).  Particles are
 * (n, 3) float positions and velocities plus int32 type (1 = liquid), level
 * l_p (radius r0 2^l_p, mass rho 8^l_p) and the deferred-coarsening counter.
 * Cell fields have ncells entries, face fields 3 * ncells (a cell's three
 * low faces).  acc holds 8 * ncells int64 fixed-point P2G accumulators
 * (cell mass, weight sum, face mass x3, face momentum x3, channel-major),
 * all zero before pf_mr_p2g; pf_mr_finish clears them again.  The scales
 * are powers of two: value = acc / scale. */
typedef struct pf_mr_params {
    float r0;                 /* base particle radius, the finest spacing */
    float rho_l, rho_g;       /* phase densities */
    float rho_min;            /* eta_phi rho_g rho~_0: phase-field floor */
    float alpha_phi;          /* eq. (7) stiffness */
    float rho_0l;             /* rho~_0 rho_l: eq. (7) liquid reference */
    float rho_tilde_0;        /* weight sum of a level-0 lattice cell */
    float k_div;              /* density-correction gain (0: off) */
    float gravity[3];
    float dt;
    float nu_l, nu_g;         /* per-phase viscosity that sets alpha_FLIP */
    float alpha_cap;          /* upper clamp of alpha_FLIP */
    float lo[3], hi[3];       /* advection clamp box */
    int32_t max_level;
    double mass_scale, mom_scale, w_scale;
} pf_mr_params;

typedef struct pf_mr_particles {
    float* x;
    float* v;
    int32_t* type;
    int32_t* level;
    int32_t* defer;
    int64_t n;                /* particles (pf_mr_emit's output: its capacity) */
} pf_mr_particles;

typedef struct pf_mr_fields {
    int64_t* acc;
    float* cwsum;
    float* phi;
    int32_t* is_liquid;
    float* fmass;
    float* fvel;
    float* fvel_star;
    float* fdvel;
    float* fbeta;
    float* rhs;
    float* pressure;
} pf_mr_fields;

/* Steps 1-3, the refinement map (particles.py:49-85): level_map (nbx nby nbz,
 * x-major) = max_level, then min(l_p) over the blocks each particle's box of
 * half width cfl_eps touches. */
int pf_mr_refine_map(const pf_mr_particles* ps, int32_t* level_map, int32_t nbx, int32_t nby, int32_t nbz,
                     float block_extent, float cfl_eps, int32_t max_level, pf_stream_t stream);
/* particle_max_speed (mr_kernels.py:507-510): out[0] = max |v|. */
int pf_mr_max_speed(const float* v, int64_t n, float* out, pf_stream_t stream);
/* Step 4, p2g (mr_kernels.py:48-133) with derive_mass_radius
 * (particles.py:234-248) folded in: adds into acc. */
int pf_mr_p2g(const pf_msbg_grid* g, const pf_mr_particles* ps, const pf_mr_params* prm, pf_mr_fields* f,
              pf_stream_t stream);
/* pf_mr_p2g over a cell-sorted visit order with a per-workgroup LDS cache
 * of cells (a counting sort by cell address, then the splat); the same
 * sums as pf_mr_p2g bit for bit.  scratch holds
 * pf_mr_p2g_scratch_ints(n, ncells) int32. */
int64_t pf_mr_p2g_scratch_ints(int64_t n, int32_t ncells);
int pf_mr_p2g_sorted(const pf_msbg_grid* g, const pf_mr_particles* ps, const pf_mr_params* prm, pf_mr_fields* f,
                     int32_t* scratch, pf_stream_t stream);
/* Step 5, face_velocity + cell_phase (mr_kernels.py:136-165): cwsum, phi,
 * is_liquid, fmass, fvel from acc; clears acc. */
int pf_mr_finish(const pf_msbg_grid* g, const pf_mr_params* prm, pf_mr_fields* f, pf_stream_t stream);
/* face_beta + zero_wall_velocity + the u* copy + add_gravity
 * (mr_kernels.py:168-227, mr_sim.py:233-245): fbeta, fvel_star, fvel. */
int pf_mr_face_prepare(const pf_msbg_grid* g, const pf_mr_params* prm, pf_mr_fields* f, pf_stream_t stream);
/* build_rhs (mr_kernels.py:286-313): rhs. */
int pf_mr_rhs(const pf_msbg_grid* g, const pf_mr_params* prm, pf_mr_fields* f, pf_stream_t stream);
/* project + field_difference (mr_kernels.py:316-343, 415-418): fvel, fdvel. */
int pf_mr_project(const pf_msbg_grid* g, const pf_mr_params* prm, pf_mr_fields* f, pf_stream_t stream);
/* max_divergence (mr_kernels.py:346-353): out[0] = max |div u|. */
int pf_mr_max_divergence(const pf_msbg_grid* g, const float* fvel, float* out, pf_stream_t stream);
/* per_particle_alpha_flip + g2p (particles.py:251-276, mr_kernels.py:413-430): v. */
int pf_mr_g2p(const pf_msbg_grid* g, const pf_mr_particles* ps, const pf_mr_params* prm, const pf_mr_fields* f,
              pf_stream_t stream);
/* Step 7, advect_rk3 + reflect_at_walls (mr_kernels.py:421-461): x, v. */
int pf_mr_advect(const pf_msbg_grid* g, const pf_mr_particles* ps, const pf_mr_params* prm, const float* fvel,
                 pf_stream_t stream);
/* Step 6a, interface_flag_per_block + block_distance_field
 * (mr_kernels.py:464-477, particles.py:387-419, mr_sim.py:163-182):
 * mixed[blk] = 1 where a block holds both phases, dist (nbx nby nbz) =
 * Euclidean distance in blocks to the nearest such block, at most max_dist
 * (max_dist everywhere without one); tmp holds 2 nbx nby nbz ints. */
int pf_mr_block_distance(const pf_msbg_grid* g, const int32_t* is_liquid, int32_t* mixed, int32_t* tmp, float* dist,
                         float max_dist, pf_stream_t stream);
/* Step 6b, classify_particles (particles.py:88-168) and the prefix sum of
 * resample (particles.py:331-372): lnew, nout (0 delete, 1 keep, 8 split),
 * defer updated, offsets[0..n] with offsets[n] = output count; scan_tmp
 * holds pf_mr_scan_tmp_ints(n) ints. */
int64_t pf_mr_scan_tmp_ints(int64_t n);
int pf_mr_classify(const pf_mr_particles* ps, const float* dist, int32_t nbx, int32_t nby, int32_t nbz,
                   float block_extent, float beta_bw, int32_t max_level, int32_t defer_threshold, int32_t seed,
                   int32_t* lnew, int32_t* nout, int32_t* offsets, int32_t* scan_tmp, pf_stream_t stream);
/* Step 6c, emit_particles (particles.py:171-231): survivors and split
 * children into out (out->n = its capacity, at least offsets[n]). */
int pf_mr_emit(const pf_mr_particles* in, const int32_t* lnew, const int32_t* nout, const int32_t* offsets,
               int32_t defer_threshold, float r0, int32_t seed, pf_mr_particles* out, pf_stream_t stream);

/* ---------------------------------------------------------------- adjoint
 * Reverse-mode derivative of one step with dt held fixed (no reference
 * counterpart: the reference has no adjoint).  Each entry point transposes
 * one forward stage at the forward state, which the caller rebuilds from a
 * checkpoint of (x, v).  Order of a backward step:
 *   pf_adj_g2p_advect -> pf_adj_extrapolate -> pf_adj_project ->
 *   pf_mgpcg_solve(lam = A^+ p-bar) -> pf_adj_solve_beta -> pf_adj_phase ->
 *   pf_adj_p2g.
 * Liquid tags, wall/solid masks and extrapolation masks are held at their
 * forward values; the eq. (7) clamps use one-sided subgradients. */

/* Transpose of pf_g2p_advect with the same args (px, pv: the step's input
 * state; vel, dvel, vel_liq: its grid fields; packed flags unsupported; the
 * solids push-out moves x by a constant and zeroes an inward velocity
 * component, so it passes x-bar and zeroes that component of v-bar).  Writes gx_in, gv_in ((n,3) adjoints of x, v) from
 * gx_out, gv_out (adjoints of x', v'); adds the grid adjoints into g_vel,
 * g_dvel, g_vel_liq with float atomics (the caller zeroes them). */
int pf_adj_g2p_advect(const pf_g2p_args* fwd, const float* gx_out, const float* gv_out, float* gx_in,
                      float* gv_in, float* const g_vel[3], float* const g_dvel[3],
                      float* const g_vel_liq[3], pf_stream_t stream);
/* Transpose of pf_extrapolate: adds E^T g_vel_liq into g_vel; g_vel_liq is
 * used as scratch; level (3 uint8 face arrays) is scratch. */
int pf_adj_extrapolate(const int32_t* is_liquid, float* const g_vel_liq[3], float* const g_vel[3],
                       uint8_t* const level[3], int nx, int ny, int nz, int sweeps, pf_stream_t stream);
/* Transpose of pf_project + dvel: in g_vel (adjoint of the projected
 * velocity) and g_dvel; out g_vel = adjoint of the pre-projection velocity,
 * g_dvel = adjoint of vel_star, g_beta and g_p (both overwritten). */
int pf_adj_project(const float* p, float* const beta[3], float* const g_vel[3], float* const g_dvel[3],
                   float* const g_beta[3], float* g_p, int nx, int ny, int nz, float dx, float dt,
                   pf_stream_t stream);
/* beta-bar of the solve: g_beta_f -= (lam_i - lam_j)(p_i - p_j) over the
 * interior faces, lam = A^+ p-bar (mean removed). */
int pf_adj_solve_beta(const float* p, const float* lam, float* const g_beta[3], int nx, int ny, int nz,
                      pf_stream_t stream);
/* Transpose of pf_phase_beta_forces (PF_PH_VEL_FROM_MOM | PHASE | ZERO |
 * STAR | GRAVITY | RHS, k_div = 0) with the same args at the forward state:
 * g_rhs = lam; writes g_mass, g_mom (face arrays). */
int pf_adj_phase(const pf_phase_args* fwd, const float* const g_vel[3], const float* const g_vel_star[3],
                 const float* const g_beta[3], const float* g_rhs, float* const g_mass[3],
                 float* const g_mom[3], pf_stream_t stream);
/* Transpose of pf_p2g's face sums (the cell mass carries no gradient):
 * adds into gx, gv ((n,3)). */
int pf_adj_p2g(const float* px, const float* pv, const float* pmass, const int32_t* pstate, int64_t n,
               const float* const g_mass[3], const float* const g_mom[3], int nx, int ny, int nz, float dx,
               float rp, float* gx, float* gv, pf_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* PFFLIP_H */
