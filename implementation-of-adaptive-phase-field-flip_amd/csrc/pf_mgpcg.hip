// Multigrid-preconditioned flexible CG for A p = b with
// (A p)_i = sum_f beta_f (p_i - p_nbr) on the MAC grid.
//
// Hierarchy: unsmoothed-aggregation Galerkin (coarse face = sum of 4 fine
// faces), red-black weighted Gauss-Seidel (omega = 6/7), V-cycle with
// pre (0,1) / post (1,0) colour orders.  Large levels run one kernel per
// red-black sweep out of HBM (one per half-sweep on request), levels of up
// to 32^3 cells below level 0 one kernel per smoothing phase; every level
// from `small_from` down is solved inside one workgroup with all its data
// in LDS.  The FPCG scalars live on the
// device; dot products are deterministic FP64 block reductions, and the
// host polls a "done" flag while the GPU still has queued work.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "pf_common.hpp"

namespace pf {

int fill_u32(void* dst, int64_t n, uint32_t v, hipStream_t s);

constexpr int MAXB = 1024;  // blocks of a reduction launch
constexpr int RT = 256;     // threads of element-wise / reduction kernels
// k_cg_apply and k_cg_beta: few, wide blocks, so the kernel that combines
// their partials itself reads few of them
constexpr int WT = 1024;
constexpr int WB = 256;

// Largest coarsest level whose smoother is applied as a precomputed matrix.
constexpr int kCoarseOpMax = 512;

struct CGState {
    double alpha, beta, mu, b_norm, rel, qAq, mu_x, tol;
    // <r, z - mu> of the iteration whose search direction parity is the
    // index: the k_cg_apply of iteration `slot` forms its value from the
    // previous one's, so the two never share a word
    double rz[2];
    int it, done, seq;  // seq: residual tests run in this solve
    int beta_parts;     // 1: the FPCG beta partials in part[0..5) are this iteration's to combine
    // 16-bit V-cycles: the level-0 right-hand side is read times vscale (a
    // power of two putting max|r| in [1/2, 1)) and z is read back times
    // vscale_inv, so the FP16 iterates stay in range as r shrinks
    float vscale, vscale_inv;
    unsigned ticket[8];
    double part[6 * MAXB];  // [0, 5 MAXB): dot-product partials; [5 MAXB, 6 MAXB): <q, Aq> partials
    float fpart[MAXB];
    // the coarsest level's `coarsest` red-black sweeps from zero followed by
    // mean removal, a linear map of its right-hand side: column j (the image
    // of the unit vector e_j) at [j * n, (j + 1) * n), built by pf_mg_build.
    // FP64: at 1000:1 contrast its entries are large and of both signs, and
    // an FP32 map loses the result to cancellation
    double coarse_op[kCoarseOpMax * kCoarseOpMax];
};

// vscale = 2^-e with m = f 2^e, f in [1/2, 1): m * vscale in [1/2, 1)
__device__ inline int vscale_exponent(float m) {
    int e = 0;
    if (m > 0.0f && m < 3.0e38f) {
        (void)frexpf(m, &e);
        e = max(min(e, 120), -120);
    }
    return e;
}

__device__ inline void set_vscale(CGState* st, float m) {
    const int e = vscale_exponent(m);
    st->vscale = ldexpf(1.0f, -e);
    st->vscale_inv = ldexpf(1.0f, e);
}

// The residual test (poisson.py:394-398) on m = max|r|, recorded by one
// thread: rel, done, the 16-bit input scale, then (number of tests so far) *
// 2 + (done != 0) to host_flag (pinned host memory, may be NULL) with a
// system-scope store the polling host sees without any event.  A solve that
// is already done only publishes.
__device__ inline void residual_test_record(CGState* st, float m, int* host_flag) {
    if (!st->done) {
        const double rel = (double)m / st->b_norm;
        st->rel = rel;
        if (rel < st->tol) st->done = 1;
        set_vscale(st, m);
    }
    st->seq += 1;
    if (host_flag)
        __hip_atomic_store(host_flag, st->seq * 2 + (st->done != 0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Pinned host words of a solver: the residual tests' flag, then the
// solve's result (its number, the iteration count, the residual as the bits
// of a double).
// HW_WANT: the number the next result is to carry, written by the host
// before the launch (a captured graph cannot take it as an argument).
enum HostWord { HW_FLAG = 0, HW_RESULT_NO = 1, HW_IT = 2, HW_WANT = 3, HW_REL = 4, HW_WORDS = 8 };

// The solve's result for the host: the iteration count and the residual,
// then the solve's number, into pinned host memory with system-scope stores
// (the number last, with release order), so the host spins on one word
// instead of copying the scalars back and synchronising the stream.  The
// kernels before it on the stream have completed when it runs.
__global__ void k_publish_result(const CGState* st, int* words) {
    if (threadIdx.x != 0) return;
    const int number = __hip_atomic_load(words + HW_WANT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const int it = st->it;
    const unsigned long long rel = (unsigned long long)__double_as_longlong(st->rel);
    __hip_atomic_store(words + HW_IT, it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(words + HW_REL), rel, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(words + HW_RESULT_NO, number, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The residual test folded into the V-cycle's first kernel (the zero-start
// level-0 sweep): every block takes the max of k_cg_update's partials itself
// (max is exact, so all blocks agree) and so knows whether to store and the
// 16-bit input scale; block 0 records the test.  st == NULL: no fold.
struct UpdateFold {
    CGState* st;
    int nparts;
    int* host_flag;
};

// The flexible CG's inner products after the preconditioner (k_cg_beta's
// work), folded into the V-cycle's last level-0 sweep, which has the final
// z and the residual r (its right-hand side) of its tile in registers: per
// workgroup the FP64 partials of sum z, sum r, sum Ap, <r, z>, <Ap, z> of
// its tile's cells, with z as stored (times zscale for 16-bit iterates).
struct DotFold {
    const float* Ap;
    const float* zscale;
    CGState* st;
};

static unsigned red_blocks(int64_t n) {
    const int64_t nb = cdiv(n, (int64_t)RT * 4);
    return (unsigned)(nb < 1 ? 1 : (nb > MAXB ? MAXB : nb));
}

// Combine per-block partials in the last block: every thread sums a fixed
// strided subset, then a fixed-order block reduction (deterministic).
template <int NV>
__device__ inline void combine_sums(const double* part, int nparts, double (&v)[NV], double* sh) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] += part[q * MAXB + b];
    block_sum<NV>(v, sh);
}

__device__ inline float combine_max(const float* part, int nparts, float* sh) {
    float m = 0.0f;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) m = fmaxf(m, part[b]);
    return block_max(m, sh);
}

// ================================================================ stencil
template <typename FT, typename PT>
__device__ inline void stencil_s(const Dims& d, const FT* __restrict__ bx, const FT* __restrict__ by,
                                 const FT* __restrict__ bz, const PT* p, int i, int j, int k,
                                 float& diag, float& s, float& pc) {
    const float bxl = to_f32(bx[face_index<0>(d, i, j, k)]);
    const float bxh = to_f32(bx[face_index<0>(d, i + 1, j, k)]);
    const float byl = to_f32(by[face_index<1>(d, i, j, k)]);
    const float byh = to_f32(by[face_index<1>(d, i, j + 1, k)]);
    const float bzl = to_f32(bz[face_index<2>(d, i, j, k)]);
    const float bzh = to_f32(bz[face_index<2>(d, i, j, k + 1)]);
    // compute_diag order (poisson.py:49-51)
    diag = bxl + bxh + byl + byh + bzl + bzh;
    pc = to_f32(p[cell_index(d, i, j, k)]);
    // _apply_row order, neighbours clamped at the walls (poisson.py:69-75)
    s = bxl * to_f32(p[cell_index(d, max(i - 1, 0), j, k)]);
    s += bxh * to_f32(p[cell_index(d, min(i + 1, d.nx - 1), j, k)]);
    s += byl * to_f32(p[cell_index(d, i, max(j - 1, 0), k)]);
    s += byh * to_f32(p[cell_index(d, i, min(j + 1, d.ny - 1), k)]);
    s += bzl * to_f32(p[cell_index(d, i, j, max(k - 1, 0))]);
    s += bzh * to_f32(p[cell_index(d, i, j, min(k + 1, d.nz - 1))]);
}

// A p in the difference form sum_f beta_f (p_i - p_f).  It equals
// diag*p_i - sum_f beta_f p_f (the reference's form, poisson.py:75), but in
// FP32 the subtractions of nearby values are exact, so the product carries
// no cancellation error: the reference's form limits the recursively
// updated CG residual to ~3e-6 (FP32) / ~4e-5 (16-bit preconditioner) at a
// 1000:1 coefficient jump, this form to ~1e-9 / ~5e-8.
// f = (x-low, x-high, y-low, y-high, z-low, z-high) faces, pn the matching
// neighbour values.  Every smoother and residual kernel goes through these
// helpers, so all of them round identically: the fused multiply-adds are
// written out and contraction is off, since the compiler could otherwise
// fuse either product of the first sum and choose differently per kernel.
__device__ inline void stencil_vals(float pc, const float (&f)[6], const float (&pn)[6], float& diag,
                                    float& Ap) {
#pragma clang fp contract(off)
    diag = f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
    float s = f[0] * (pc - pn[0]);
    s = fmaf(f[1], pc - pn[1], s);
    s = fmaf(f[2], pc - pn[2], s);
    s = fmaf(f[3], pc - pn[3], s);
    s = fmaf(f[4], pc - pn[4], s);
    s = fmaf(f[5], pc - pn[5], s);
    Ap = s;
}

// weighted Gauss-Seidel update of one cell (poisson.py:120-126)
__device__ inline float rb_relax(float pc, float b, float Ap, float dg, float omega) {
    return pc + omega * (b - Ap) / dg;
}
// the same update from p = 0 (every neighbour and the cell itself zero)
__device__ inline float rb_first(float b, float dg, float omega) { return omega * b / dg; }

template <typename ST>
__device__ inline float round_st(float v) { return to_f32(from_f32<ST>(v)); }

template <typename FT, typename PT>
__device__ inline void stencil(const Dims& d, const FT* __restrict__ bx, const FT* __restrict__ by,
                               const FT* __restrict__ bz, const PT* p, int i, int j, int k,
                               float& diag, float& Ap, float& pc) {
    const float f[6] = {to_f32(bx[face_index<0>(d, i, j, k)]), to_f32(bx[face_index<0>(d, i + 1, j, k)]),
                        to_f32(by[face_index<1>(d, i, j, k)]), to_f32(by[face_index<1>(d, i, j + 1, k)]),
                        to_f32(bz[face_index<2>(d, i, j, k)]), to_f32(bz[face_index<2>(d, i, j, k + 1)])};
    pc = to_f32(p[cell_index(d, i, j, k)]);
    const float pn[6] = {to_f32(p[cell_index(d, max(i - 1, 0), j, k)]),
                         to_f32(p[cell_index(d, min(i + 1, d.nx - 1), j, k)]),
                         to_f32(p[cell_index(d, i, max(j - 1, 0), k)]),
                         to_f32(p[cell_index(d, i, min(j + 1, d.ny - 1), k)]),
                         to_f32(p[cell_index(d, i, j, max(k - 1, 0))]),
                         to_f32(p[cell_index(d, i, j, min(k + 1, d.nz - 1))])};
    stencil_vals(pc, f, pn, diag, Ap);
}

// ================================================================ hierarchy build
// Coarse face = sum of the 4 fine faces it covers (poisson.py:130-163).
__global__ __launch_bounds__(RT) void k_galerkin(Dims df, Dims dc, const float* __restrict__ fx,
                                                 const float* __restrict__ fy,
                                                 const float* __restrict__ fz, float* __restrict__ cx,
                                                 float* __restrict__ cy, float* __restrict__ cz) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ez = dc.nz + 1, ey = dc.ny + 1;
    if (t >= (int64_t)(dc.nx + 1) * ey * ez) return;
    const int K = (int)(t % ez), J = (int)((t / ez) % ey), I = (int)(t / ((int64_t)ez * ey));
    const int i0 = 2 * I, j0 = 2 * J, k0 = 2 * K;
    if (J < dc.ny && K < dc.nz) {
        float acc = 0.f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const int jj = j0 + a, kk = k0 + b;
                if (i0 <= df.nx && jj < df.ny && kk < df.nz) acc += fx[face_index<0>(df, i0, jj, kk)];
            }
        cx[face_index<0>(dc, I, J, K)] = acc;
    }
    if (I < dc.nx && K < dc.nz) {
        float acc = 0.f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const int ii = i0 + a, kk = k0 + b;
                if (ii < df.nx && j0 <= df.ny && kk < df.nz) acc += fy[face_index<1>(df, ii, j0, kk)];
            }
        cy[face_index<1>(dc, I, J, K)] = acc;
    }
    if (I < dc.nx && J < dc.ny) {
        float acc = 0.f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const int ii = i0 + a, jj = j0 + b;
                if (ii < df.nx && jj < df.ny && k0 <= df.nz) acc += fz[face_index<2>(df, ii, jj, k0)];
            }
        cz[face_index<2>(dc, I, J, K)] = acc;
    }
}

// diag (poisson.py:40-51) and, for levels run out of HBM in reduced
// precision, the 16-bit copies of the face coefficients.
template <typename ST>
__global__ __launch_bounds__(RT) void k_diag_copy(Dims d, const float* __restrict__ bx,
                                                  const float* __restrict__ by,
                                                  const float* __restrict__ bz, float* diag, ST* hx,
                                                  ST* hy, ST* hz) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ez = d.nz + 1, ey = d.ny + 1;
    if (t >= (int64_t)(d.nx + 1) * ey * ez) return;
    const int k = (int)(t % ez), j = (int)((t / ez) % ey), i = (int)(t / ((int64_t)ez * ey));
    const bool ci = i < d.nx, cj = j < d.ny, ck = k < d.nz;
    if (hx) {
        if (cj && ck) hx[face_index<0>(d, i, j, k)] = from_f32<ST>(bx[face_index<0>(d, i, j, k)]);
        if (ci && ck) hy[face_index<1>(d, i, j, k)] = from_f32<ST>(by[face_index<1>(d, i, j, k)]);
        if (ci && cj) hz[face_index<2>(d, i, j, k)] = from_f32<ST>(bz[face_index<2>(d, i, j, k)]);
    }
    if (diag && ci && cj && ck)
        diag[cell_index(d, i, j, k)] = bx[face_index<0>(d, i, j, k)] + bx[face_index<0>(d, i + 1, j, k)] +
                                       by[face_index<1>(d, i, j, k)] + by[face_index<1>(d, i, j + 1, k)] +
                                       bz[face_index<2>(d, i, j, k)] + bz[face_index<2>(d, i, j, k + 1)];
}

// ================================================================ HBM smoother / transfers
// Red-black half-sweep (poisson.py:105-126).  One thread per cell of the
// active colour.  ZINIT: p is known to be zero (a freshly cleared level),
// so res = b and every cell of the other colour is written as 0.
template <typename ST, bool ZINIT>
__global__ __launch_bounds__(RT) void k_rbgs(Dims d, const ST* __restrict__ bx,
                                             const ST* __restrict__ by, const ST* __restrict__ bz,
                                             ST* p, const float* __restrict__ b, int colour,
                                             float omega, const int* __restrict__ done,
                                             const float* __restrict__ bscale) {
    // the flag is only consulted at the stores, so its load does not delay
    // the stencil loads (a converged solve's extra iteration stores nothing)
    const bool skip = done && *done;
    const float bs = bscale ? *bscale : 1.0f;
    const int nkp = (d.nz + 1) >> 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)d.nx * d.ny * nkp) return;
    const int kp = (int)(t % nkp);
    const int row = (int)(t / nkp);
    const int j = row % d.ny, i = row / d.ny;
    if (ZINIT) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int k = 2 * kp + q;
            if (k >= d.nz) break;
            float v = 0.0f;
            if (((i + j + k) & 1) == colour) {
                const float dg = to_f32(bx[face_index<0>(d, i, j, k)]) + to_f32(bx[face_index<0>(d, i + 1, j, k)]) +
                                 to_f32(by[face_index<1>(d, i, j, k)]) + to_f32(by[face_index<1>(d, i, j + 1, k)]) +
                                 to_f32(bz[face_index<2>(d, i, j, k)]) + to_f32(bz[face_index<2>(d, i, j, k + 1)]);
                if (dg > 0.0f) v = rb_first(b[cell_index(d, i, j, k)] * bs, dg, omega);
            }
            if (!skip) p[cell_index(d, i, j, k)] = from_f32<ST>(v);
        }
    } else {
        const int k = 2 * kp + ((i + j + colour) & 1);
        if (k >= d.nz) return;
        float dg, Ap, pc;
        stencil(d, bx, by, bz, p, i, j, k, dg, Ap, pc);
        if (dg <= 0.0f) return;
        const int c = cell_index(d, i, j, k);
        if (!skip) p[c] = from_f32<ST>(rb_relax(pc, b[c] * bs, Ap, dg, omega));
    }
}

// Pair loads/stores: cells (i,j,k) and (i,j,k+1) with k even, adjacent in
// memory and 2-element aligned whenever the row length is even.
template <typename ST>
__device__ inline void ld_pair(const ST* p, int idx, float& a, float& b);
template <>
__device__ inline void ld_pair<float>(const float* p, int idx, float& a, float& b) {
    const float2 v = *reinterpret_cast<const float2*>(p + idx);
    a = v.x;
    b = v.y;
}
template <typename ST>
__device__ inline void ld_pair(const ST* p, int idx, float& a, float& b) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(p + idx);
    a = to_f32(ST{(uint16_t)(w & 0xFFFFu)});
    b = to_f32(ST{(uint16_t)(w >> 16)});
}
template <typename ST>
__device__ inline void st_pair(ST* p, int idx, float a, float b);
template <>
__device__ inline void st_pair<float>(float* p, int idx, float a, float b) {
    *reinterpret_cast<float2*>(p + idx) = make_float2(a, b);
}
template <typename ST>
__device__ inline void st_pair(ST* p, int idx, float a, float b) {
    const uint32_t w = (uint32_t)from_f32<ST>(a).u | ((uint32_t)from_f32<ST>(b).u << 16);
    *reinterpret_cast<uint32_t*>(p + idx) = w;
}

// One full red-black sweep (colour c1, then the other colour) in one launch,
// bitwise equal to two in-place half-sweeps.  A workgroup owns a TI x TJ x TK
// tile.  It stages the iterate of the tile plus a 2-cell halo in LDS
// (coordinates clamped into the grid, which reproduces the reference's
// clamped neighbour indices: a clamped neighbour is the cell itself), updates
// the first colour on the tile and on its 1-cell halo shell (the shell
// updates duplicate work of the neighbouring tiles), then the second colour
// on the tile, and writes the whole tile to `pout`.  Reading `pin` and
// writing `pout` (ping-pong) keeps neighbouring tiles' halo reads race-free.
// Each thread owns pairs of cells (k, k+1) of the tile and keeps the pair's
// faces in registers between the two colours, so faces, b and p are read
// from memory once per sweep.  Every global load (staging, the tile's faces
// and b, the shell cells' faces and b) is issued before the first wait, so
// a workgroup is exposed to memory latency once.  Needs nz even.
// ZINIT: the input is zero (no staging loads).
template <int TI, int TJ, int TK, int NTH>
struct RbTile {
    static constexpr int NT = NTH;
    static constexpr int PI = TI + 4, PJ = TJ + 4, PK = TK + 4;
    static constexpr int NP = PI * PJ * PK;
    static constexpr int NPR = PK / 2;                    // staged pairs per row
    static constexpr int NSP = PI * PJ * NPR;             // staged pairs
    static constexpr int SPT = (NSP + NT - 1) / NT;       // staged pairs per thread
    static constexpr int PAIRS = TI * TJ * TK / 2;
    static constexpr int PPT = PAIRS / NT;
    // first-colour shell cells: i-planes and j-planes hold rows along k of
    // TK+2 cells (half of each row has colour c1); k-planes rows along j
    static constexpr int RK = (TK + 2) / 2;
    static constexpr int S1 = 2 * (TJ + 2) * RK, S2 = 2 * TI * RK, S3 = 2 * TI * (TJ / 2);
    static constexpr int NS = S1 + S2 + S3;
    static constexpr int SHT = (NS + NT - 1) / NT;
    static_assert(PAIRS % NT == 0 && TK % 2 == 0 && TJ % 2 == 0 && PPT >= 1, "tile shape");
};

// (4 waves per SIMD: a 128^3 level's 512 workgroups of 8x8x64 tiles are all
// resident at once, so the sweep costs one memory round trip, not two)
// PROL: the sweep's input is pin plus the piecewise-constant prolongation of
// the coarse iterate pc (poisson.py:178-182), formed in LDS while staging
// (rounded to ST as the separate prolongation stores it): the first
// post-sweep of a level without a separate prolongation pass.
// DOTS: the CG's inner products of the sweep's output (DotFold), with Ap
// staged into LDS alongside the iterate and the sums formed after the
// stores, when the face registers are free.
template <typename ST, bool ZINIT, int TI, int TJ, int TK, int NTH = 256, bool PROL = false, typename CT = float,
          bool DOTS = false>
__global__ __launch_bounds__(NTH, 4) void k_rbgs_full(Dims d, const ST* __restrict__ bx,
                                                   const ST* __restrict__ by, const ST* __restrict__ bz,
                                                   const ST* __restrict__ pin, ST* __restrict__ pout,
                                                   const float* __restrict__ b, int c1, float omega,
                                                   const int* __restrict__ done,
                                                   const float* __restrict__ bscale,
                                                   const CT* __restrict__ pcoarse, Dims dc, UpdateFold uf,
                                                   DotFold df) {
    using T = RbTile<TI, TJ, TK, NTH>;
    __shared__ __attribute__((aligned(16))) float P[T::NP];
    __shared__ float s_fm[NTH / 64];
    const int tid = threadIdx.x;
    // the residual test's inputs, loaded with everything else
    const bool fold = uf.st != nullptr;
    int was_done = 0;
    float fm = 0.0f;
    if (fold) {
        was_done = uf.st->done;
        for (int e = tid; e < uf.nparts; e += T::NT) fm = fmaxf(fm, uf.st->fpart[e]);
    }
    const int ntk = (d.nz + TK - 1) / TK, ntj = (d.ny + TJ - 1) / TJ;
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = (tile / (ntk * ntj)) * TI, j0 = ((tile / ntk) % ntj) * TJ, k0 = (tile % ntk) * TK;
    constexpr int SI = T::PJ * T::PK, SJ = T::PK;
    auto lidx = [&](int i, int j, int k) { return ((i - i0 + 2) * T::PJ + (j - j0 + 2)) * T::PK + (k - k0 + 2); };

    // (1) staging: the iterate of the tile + 2-cell halo goes straight from
    // memory into LDS (global_load_lds: lane l of a wave writes its dword to
    // the wave's LDS base + 4 l, no registers).  FP32: one element per lane
    // from clamped coordinates.  16-bit: raw pairs (k, k+1) from the nearest
    // in-grid pair into R, converted into P after the barrier.
    constexpr bool F32 = sizeof(ST) == 4;
    __shared__ uint32_t R[F32 ? 1 : T::NSP];
    if (!ZINIT) {
        if constexpr (F32) {
            for (int e0 = 0; e0 < T::NP; e0 += T::NT) {
                const int e = e0 + tid;
                if (e < T::NP) {
                    const int kk = e % T::PK, jj = (e / T::PK) % T::PJ, ii = e / SI;
                    const int gi = min(max(i0 - 2 + ii, 0), d.nx - 1);
                    const int gj = min(max(j0 - 2 + jj, 0), d.ny - 1);
                    const int gk = min(max(k0 - 2 + kk, 0), d.nz - 1);
                    __builtin_amdgcn_global_load_lds(
                        (const void*)(pin + cell_index(d, gi, gj, gk)),
                        (__attribute__((address_space(3))) void*)(P + e0 + (tid & ~63)), 4, 0, 0);
                }
            }
        } else {
            for (int e0 = 0; e0 < T::NSP; e0 += T::NT) {
                const int e = e0 + tid;
                if (e < T::NSP) {
                    const int q = e % T::NPR, row = e / T::NPR;
                    const int gi = min(max(i0 - 2 + row / T::PJ, 0), d.nx - 1);
                    const int gj = min(max(j0 - 2 + row % T::PJ, 0), d.ny - 1);
                    const int kc = min(max(k0 - 2 + 2 * q, 0), d.nz - 2);
                    __builtin_amdgcn_global_load_lds(
                        (const void*)(pin + cell_index(d, gi, gj, kc)),
                        (__attribute__((address_space(3))) void*)(R + e0 + (tid & ~63)), 4, 0, 0);
                }
            }
        }
    }

    // (1b) PROL: the coarse iterate under the staged region, parents
    // [i0/2 - 1, i0/2 + TI/2 + 1) per axis (coordinates clamped)
    constexpr int CI = TI / 2 + 2, CJ = TJ / 2 + 2, CK = TK / 2 + 2, CN = CI * CJ * CK;
    constexpr int CPT = PROL ? (CN + T::NT - 1) / T::NT : 1;
    float cv[CPT];
    if constexpr (PROL) {
#pragma unroll
        for (int q = 0; q < CPT; ++q) {
            const int e = tid + q * T::NT;
            cv[q] = 0.0f;
            if (e < CN) {
                const int ck = e % CK, cj = (e / CK) % CJ, ci = e / (CK * CJ);
                const int gi = min(max(i0 / 2 - 1 + ci, 0), dc.nx - 1);
                const int gj = min(max(j0 / 2 - 1 + cj, 0), dc.ny - 1);
                const int gk = min(max(k0 / 2 - 1 + ck, 0), dc.nz - 1);
                cv[q] = to_f32(pcoarse[cell_index(dc, gi, gj, gk)]);
            }
        }
    }

    // (2) the tile's pairs: faces and b into registers
    float f[T::PPT][11], bb[T::PPT][2];
    int li[T::PPT];
    bool ok[T::PPT];
#pragma unroll
    for (int q = 0; q < T::PPT; ++q) {
        const int pq = tid + q * T::NT;
        const int kp = pq % (TK / 2), jl = (pq / (TK / 2)) % TJ, il = pq / ((TK / 2) * TJ);
        const int i = i0 + il, j = j0 + jl, k = k0 + 2 * kp;
        ok[q] = i < d.nx && j < d.ny && k < d.nz;
        li[q] = ((il + 2) * T::PJ + (jl + 2)) * T::PK + (2 * kp + 2);
        if (ok[q]) {
            ld_pair(bx, face_index<0>(d, i, j, k), f[q][0], f[q][1]);      // x-low of k, k+1
            ld_pair(bx, face_index<0>(d, i + 1, j, k), f[q][2], f[q][3]);  // x-high
            ld_pair(by, face_index<1>(d, i, j, k), f[q][4], f[q][5]);      // y-low
            ld_pair(by, face_index<1>(d, i, j + 1, k), f[q][6], f[q][7]);  // y-high
            const int z = face_index<2>(d, i, j, k);
            f[q][8] = to_f32(bz[z]);
            f[q][9] = to_f32(bz[z + 1]);
            f[q][10] = to_f32(bz[z + 2]);
            ld_pair(b, cell_index(d, i, j, k), bb[q][0], bb[q][1]);
        }
    }
    // (1c) DOTS: Ap of the tile into LDS, cell (il, jl, kl) at its tile-linear
    // index (every lane loads: coordinates clamped into the grid)
    constexpr int TC = TI * TJ * TK;
    __shared__ float A[DOTS ? TC : 1];
    if constexpr (DOTS) {
        static_assert(TC % T::NT == 0, "whole passes");
        for (int e0 = 0; e0 < TC; e0 += T::NT) {
            const int e = e0 + tid;
            const int kl = e % TK, jl = (e / TK) % TJ, il = e / (TK * TJ);
            const int gi = min(i0 + il, d.nx - 1), gj = min(j0 + jl, d.ny - 1), gk = min(k0 + kl, d.nz - 1);
            __builtin_amdgcn_global_load_lds((const void*)(df.Ap + cell_index(d, gi, gj, gk)),
                                             (__attribute__((address_space(3))) void*)(A + e0 + (tid & ~63)), 4,
                                             0, 0);
        }
    }

    // (3) first-colour cells of the 1-cell halo shell that lie in the grid
    float sf[T::SHT][6], sb[T::SHT];
    int sc[T::SHT];
    bool sok[T::SHT];
#pragma unroll
    for (int r = 0; r < T::SHT; ++r) {
        int e = tid + r * T::NT, i, j, k;
        if (e < T::S1) {  // i = i0-1 and i0+TI, all j of the shell, rows along k
            i = e < T::S1 / 2 ? i0 - 1 : i0 + TI;
            e %= T::S1 / 2;
            j = j0 - 1 + e / T::RK;
            const int m = e % T::RK, base = k0 - 1;
            k = base + 2 * m + ((c1 - (i + j + base)) & 1);
        } else if ((e -= T::S1) < T::S2) {  // j = j0-1 and j0+TJ, i in the tile
            j = e < T::S2 / 2 ? j0 - 1 : j0 + TJ;
            e %= T::S2 / 2;
            i = i0 + e / T::RK;
            const int m = e % T::RK, base = k0 - 1;
            k = base + 2 * m + ((c1 - (i + j + base)) & 1);
        } else {  // k = k0-1 and k0+TK, i and j in the tile, rows along j
            e -= T::S2;
            k = e < T::S3 / 2 ? k0 - 1 : k0 + TK;
            e %= T::S3 / 2;
            i = i0 + e / (TJ / 2);
            const int m = e % (TJ / 2);
            j = j0 + 2 * m + ((c1 - (i + j0 + k)) & 1);
        }
        sok[r] = (tid + r * T::NT) < T::NS && i >= 0 && j >= 0 && k >= 0 && i < d.nx && j < d.ny && k < d.nz;
        sc[r] = lidx(i, j, k);
        if (sok[r]) {
            sf[r][0] = to_f32(bx[face_index<0>(d, i, j, k)]);
            sf[r][1] = to_f32(bx[face_index<0>(d, i + 1, j, k)]);
            sf[r][2] = to_f32(by[face_index<1>(d, i, j, k)]);
            sf[r][3] = to_f32(by[face_index<1>(d, i, j + 1, k)]);
            sf[r][4] = to_f32(bz[face_index<2>(d, i, j, k)]);
            sf[r][5] = to_f32(bz[face_index<2>(d, i, j, k + 1)]);
            sb[r] = b[cell_index(d, i, j, k)];
        }
    }
    bool skip = !fold && done && *done;
    if (fold) {
        fm = wave_max(fm);
        if ((tid & 63) == 0) s_fm[tid >> 6] = fm;
    }

    if (ZINIT) {
        for (int e = tid; e < T::NP; e += T::NT) P[e] = 0.0f;
    } else {
        __builtin_amdgcn_s_waitcnt(0);  // the LDS-destined loads have landed
    }
    __syncthreads();
    float bs = 1.0f;
    if (fold) {
        float mm = s_fm[0];
#pragma unroll
        for (int w = 1; w < NTH / 64; ++w) mm = fmaxf(mm, s_fm[w]);
        const bool conv = (double)mm / uf.st->b_norm < uf.st->tol;
        skip = was_done || conv;
        bs = ldexpf(1.0f, -vscale_exponent(mm));
        if (blockIdx.x == 0 && tid == 0) residual_test_record(uf.st, mm, uf.host_flag);
    } else if (bscale) {
        bs = *bscale;
    }
    // a finished solve's queued iterations stop here: after the loads (the
    // flag is not waited for before issuing them) but before any arithmetic
    // (every load has landed; skip is the same in the whole workgroup)
    if (skip) return;
    if (bscale) {
#pragma unroll
        for (int q = 0; q < T::PPT; ++q) {
            bb[q][0] *= bs;
            bb[q][1] *= bs;
        }
#pragma unroll
        for (int r = 0; r < T::SHT; ++r) sb[r] *= bs;
    }
    if constexpr (!F32) {
        if (!ZINIT) {
            // raw pairs -> floats (LDS index 2 * pair index); past either end
            // of a row both entries clamp to the end cell
            for (int e = tid; e < T::NSP; e += T::NT) {
                const uint32_t w = R[e];
                float x = to_f32(ST{(uint16_t)(w & 0xFFFFu)}), y = to_f32(ST{(uint16_t)(w >> 16)});
                const int kk = k0 - 2 + 2 * (e % T::NPR);
                if (kk < 0) y = x;
                if (kk > d.nz - 2) x = y;
                P[2 * e] = x;
                P[2 * e + 1] = y;
            }
            __syncthreads();
        }
    }
    if constexpr (PROL) {
        __shared__ float C[PROL ? CN : 1];
#pragma unroll
        for (int q = 0; q < CPT; ++q)
            if (tid + q * T::NT < CN) C[tid + q * T::NT] = cv[q];
        __syncthreads();
        // staged pairs (k, k+1) with k even share their parent, clamping
        // included (k0 is even, so a clamped pair clamps to one cell)
        for (int e2 = tid; e2 < T::NP / 2; e2 += T::NT) {
            const int m = e2 % (T::PK / 2), row = e2 / (T::PK / 2);
            const int jj = row % T::PJ, ii = row / T::PJ;
            const int gi = min(max(i0 - 2 + ii, 0), d.nx - 1);
            const int gj = min(max(j0 - 2 + jj, 0), d.ny - 1);
            const int gk = min(max(k0 - 2 + 2 * m, 0), d.nz - 1);
            const int ci = (gi >> 1) - (i0 / 2 - 1), cj = (gj >> 1) - (j0 / 2 - 1), ck = (gk >> 1) - (k0 / 2 - 1);
            const float c = C[(ci * CJ + cj) * CK + ck];
            float2 v = *reinterpret_cast<float2*>(P + 2 * e2);
            v.x = round_st<ST>(v.x + c);
            v.y = round_st<ST>(v.y + c);
            *reinterpret_cast<float2*>(P + 2 * e2) = v;
        }
        __syncthreads();
    }

    // first colour on the tile
#pragma unroll
    for (int q = 0; q < T::PPT; ++q) {
        if (!ok[q]) continue;
        const int pq = tid + q * T::NT;
        const int jl = (pq / (TK / 2)) % TJ, il = pq / ((TK / 2) * TJ);
        const int h = ((i0 + il + j0 + jl) & 1) == c1 ? 0 : 1;  // which cell of the pair has colour c1
        const int c = li[q] + h;
        const float fc[6] = {f[q][h], f[q][2 + h], f[q][4 + h], f[q][6 + h], f[q][8 + h], f[q][9 + h]};
        float v;
        if (ZINIT) {
            const float dg = fc[0] + fc[1] + fc[2] + fc[3] + fc[4] + fc[5];
            v = dg > 0.0f ? rb_first(bb[q][h], dg, omega) : 0.0f;
        } else {
            const float pc = P[c];
            const float pn[6] = {P[c - SI], P[c + SI], P[c - SJ], P[c + SJ], P[c - 1], P[c + 1]};
            float dg, Ap;
            stencil_vals(pc, fc, pn, dg, Ap);
            v = dg > 0.0f ? rb_relax(pc, bb[q][h], Ap, dg, omega) : pc;
        }
        P[c] = round_st<ST>(v);
    }
    // first colour on the halo shell
#pragma unroll
    for (int r = 0; r < T::SHT; ++r) {
        if (!sok[r]) continue;
        const int c = sc[r];
        const float fc[6] = {sf[r][0], sf[r][1], sf[r][2], sf[r][3], sf[r][4], sf[r][5]};
        float v;
        if (ZINIT) {
            const float dg = fc[0] + fc[1] + fc[2] + fc[3] + fc[4] + fc[5];
            v = dg > 0.0f ? rb_first(sb[r], dg, omega) : 0.0f;
        } else {
            const float pc = P[c];
            const float pn[6] = {P[c - SI], P[c + SI], P[c - SJ], P[c + SJ], P[c - 1], P[c + 1]};
            float dg, Ap;
            stencil_vals(pc, fc, pn, dg, Ap);
            v = dg > 0.0f ? rb_relax(pc, sb[r], Ap, dg, omega) : pc;
        }
        P[c] = round_st<ST>(v);
    }
    __syncthreads();

    // second colour on the tile, then the pair goes out
#pragma unroll
    for (int q = 0; q < T::PPT; ++q) {
        if (!ok[q]) continue;
        const int pq = tid + q * T::NT;
        const int kp = pq % (TK / 2), jl = (pq / (TK / 2)) % TJ, il = pq / ((TK / 2) * TJ);
        const int h = ((i0 + il + j0 + jl) & 1) == c1 ? 1 : 0;  // the cell of the other colour
        const int c = li[q] + h;
        const float fc[6] = {f[q][h], f[q][2 + h], f[q][4 + h], f[q][6 + h], f[q][8 + h], f[q][9 + h]};
        const float pc = P[c];
        const float pn[6] = {P[c - SI], P[c + SI], P[c - SJ], P[c + SJ], P[c - 1], P[c + 1]};
        float dg, Ap;
        stencil_vals(pc, fc, pn, dg, Ap);
        const float v = dg > 0.0f ? rb_relax(pc, bb[q][h], Ap, dg, omega) : pc;
        const float other = P[li[q] + (1 - h)];
        if (!skip)
            st_pair(pout, cell_index(d, i0 + il, j0 + jl, k0 + 2 * kp), h ? other : v, h ? v : other);
        // DOTS: the cell's value as stored (no other thread reads a
        // second-colour cell in this half-sweep)
        if (DOTS) P[c] = round_st<ST>(v);
    }
    if constexpr (DOTS) {
        // k_cg_beta's sums over the tile: z as stored times zscale (its
        // zval), r = b unscaled (b * bs * zscale, powers of two), Ap
        const float zsc = df.zscale ? *df.zscale : 1.0f;
        // (A landed before the first barrier; a pair's two cells were
        // written by this thread)
        double dsum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < T::PPT; ++q) {
            if (!ok[q]) continue;
            const int pq = tid + q * T::NT;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const double zz = (double)(P[li[q] + e] * zsc), rr = (double)(bb[q][e] * zsc);
                const double ap = (double)A[2 * pq + e];
                dsum[0] += zz;
                dsum[1] += rr;
                dsum[2] += ap;
                dsum[3] += rr * zz;
                dsum[4] += ap * zz;
            }
        }
        __shared__ double s_dot[5 * 16];
        block_sum<5>(dsum, s_dot);
        if (tid == 0 && !skip) {
#pragma unroll
            for (int q = 0; q < 5; ++q) df.st->part[q * MAXB + blockIdx.x] = dsum[q];
            if (blockIdx.x == 0) df.st->beta_parts = 1;  // read by the next k_cg_apply
        }
    }
}

// residual (poisson.py:91-102) fused with R = P^T (poisson.py:166-175)
template <typename ST>
__global__ __launch_bounds__(RT) void k_resid_restrict(Dims d, Dims dc, const ST* __restrict__ bx,
                                                       const ST* __restrict__ by,
                                                       const ST* __restrict__ bz,
                                                       const ST* __restrict__ p,
                                                       const float* __restrict__ b,
                                                       float* __restrict__ bc,
                                                       const int* __restrict__ done,
                                                       const float* __restrict__ bscale) {
    const bool skip = done && *done;
    const float bs = bscale ? *bscale : 1.0f;
    // 32-bit index arithmetic (levels hold fewer than 2^31 cells)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dc.nx * dc.ny * dc.nz) return;
    const int K = t % dc.nz, J = (t / dc.nz) % dc.ny, I = t / (dc.nz * dc.ny);
    // The 2x2x2 children share their inputs: each row (i, j) of the block
    // and its four side rows are read as (k0, k0+1) pairs (k0 even and the
    // fine rows even, so a pair is aligned; the z-faces' rows are odd and
    // read singly), 48 loads instead of 8 separate stencils' 112.  The
    // neighbours are clamped at the walls as in stencil().
    const int i0 = 2 * I, j0 = 2 * J, k0 = 2 * K;
    const int im = max(i0 - 1, 0), ip = min(i0 + 2, d.nx - 1);
    const int jm = max(j0 - 1, 0), jp = min(j0 + 2, d.ny - 1);
    const int km = max(k0 - 1, 0), kp = min(k0 + 2, d.nz - 1);
    // pc[a][bb][c]: the block; px[0|1][bb][c]: rows im / ip; py[a][0|1][c]:
    // rows jm / jp; pz[a][bb][0|1]: k0-1 / k0+2 of the block's rows
    float pc[2][2][2], px[2][2][2], py[2][2][2], pz[2][2][2];
    float fx[3][2][2], fy[2][3][2], fz[2][2][3], bv[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
            const int i = i0 + a, j = j0 + bb;
            ld_pair(p, cell_index(d, i, j, k0), pc[a][bb][0], pc[a][bb][1]);
            ld_pair(p, cell_index(d, a ? ip : im, j, k0), px[a][bb][0], px[a][bb][1]);
            ld_pair(p, cell_index(d, i, bb ? jp : jm, k0), py[a][bb][0], py[a][bb][1]);
            pz[a][bb][0] = to_f32(p[cell_index(d, i, j, km)]);
            pz[a][bb][1] = to_f32(p[cell_index(d, i, j, kp)]);
            ld_pair(b, cell_index(d, i, j, k0), bv[a][bb][0], bv[a][bb][1]);
            const int z = face_index<2>(d, i, j, k0);
            fz[a][bb][0] = to_f32(bz[z]);
            fz[a][bb][1] = to_f32(bz[z + 1]);
            fz[a][bb][2] = to_f32(bz[z + 2]);
        }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
            ld_pair(bx, face_index<0>(d, i0 + a, j0 + bb, k0), fx[a][bb][0], fx[a][bb][1]);
            ld_pair(by, face_index<1>(d, i0 + bb, j0 + a, k0), fy[bb][a][0], fy[bb][a][1]);
        }
    float acc = 0.0f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const float f[6] = {fx[a][bb][c], fx[a + 1][bb][c], fy[a][bb][c],
                                    fy[a][bb + 1][c], fz[a][bb][c], fz[a][bb][c + 1]};
                // -x, +x, -y, +y, -z, +z (the sibling inside the block, or
                // the side row)
                const float pn[6] = {a ? pc[0][bb][c] : px[0][bb][c], a ? px[1][bb][c] : pc[1][bb][c],
                                     bb ? pc[a][0][c] : py[a][0][c], bb ? py[a][1][c] : pc[a][1][c],
                                     c ? pc[a][bb][0] : pz[a][bb][0], c ? pz[a][bb][1] : pc[a][bb][1]};
                float dg, Ap;
                stencil_vals(pc[a][bb][c], f, pn, dg, Ap);
                acc += bv[a][bb][c] * bs - Ap;
            }
    if (!skip) bc[t] = acc;
}

// P: piecewise-constant injection added onto the fine level (poisson.py:178-182)
// (pin and pout may be the same array)
template <typename ST, typename CT>
__global__ __launch_bounds__(RT) void k_prolong(Dims d, Dims dc, const ST* pin, ST* pout,
                                                const CT* __restrict__ pc, const int* __restrict__ done) {
    const bool skip = done && *done;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)d.nx * d.ny * d.nz) return;
    const int k = (int)(t % d.nz), j = (int)((t / d.nz) % d.ny), i = (int)(t / ((int64_t)d.nz * d.ny));
    const float v = to_f32(pin[t]) + to_f32(pc[cell_index(dc, i / 2, j / 2, k / 2)]);
    if (!skip) pout[t] = from_f32<ST>(v);
}

// ================================================================ fused small levels
// A small HBM level's whole pre-smoothing phase (sweeps from zero, colour
// order (0,1), then the residual restricted to the coarse level) or whole
// post-smoothing phase (input pin plus the prolongation of the coarse
// iterate, colour order (1,0)) in one launch: on a level of a few ten
// thousand cells each per-sweep launch is bound by launch latency, not by
// its bytes.  A workgroup owns a TI x TJ x TK tile and holds the tile plus
// an H-cell halo (two cells per sweep) of faces, right-hand side and
// iterate in LDS.  Each half-sweep runs on the region the tile's result
// still depends on (margin m of the box; it shrinks by one cell per
// half-sweep), so the halo updates duplicate the neighbouring tiles' work
// and no workgroup waits for another.  Every update goes through the
// helpers of the per-sweep kernels in their operation order, so the
// results are bitwise those of the per-sweep path.  A neighbour outside the
// grid is the cell itself (the reference's clamped indices).
template <int TI, int TJ, int TK, int NSW>
struct LvBox {
    static constexpr int NT = 1024;
    static constexpr int H = 2 * NSW;
    static constexpr int BI = TI + 2 * H, BJ = TJ + 2 * H, BK = TK + 2 * H;
    // one extra layer per axis holds the high faces of the outermost cells;
    // rows of BK + 2 (even) entries: with odd strides every cell of one
    // colour would sit at an address of one parity, so a half-sweep's
    // accesses would use half the LDS banks
    static constexpr int SJ = BK + 2, SI = (BJ + 1) * SJ, N = (BI + 1) * SI;
    static constexpr int NPT = (N + NT - 1) / NT;
    static_assert(TI % 2 == 0 && TJ % 2 == 0 && TK % 2 == 0, "tiles restrict to whole coarse cells");
};

// One half-sweep of colour col on the cells at margin >= M of the box
// (origin o, which is even on every axis).  ZI: the iterate is zero.
// A neighbour outside the grid is read from its ghost entry, the box entry
// beyond the wall, which every update of the cell sets to the cell's value
// (the reference's clamped index is the cell itself), so the six
// neighbour reads take no branches.
template <typename ST, typename BX, int M, bool ZI>
__device__ inline void lv_half_sweep(float* P, const float* B, const float* FX, const float* FY,
                                     const float* FZ, const Dims& d, int ox, int oy, int oz, int col,
                                     float omega) {
    constexpr int RJ = BX::BJ - 2 * M, RK2 = (BX::BK - 2 * M) / 2;
    constexpr int NPAIR = (BX::BI - 2 * M) * RJ * RK2;
    constexpr int SI = BX::SI, SJ = BX::SJ;
    for (int e = threadIdx.x; e < NPAIR; e += BX::NT) {
        const int kp = e % RK2, y = M + (e / RK2) % RJ, x = M + e / (RK2 * RJ);
        const int z = M + 2 * kp + ((col - (x + y + M)) & 1);
        const int gi = ox + x, gj = oy + y, gk = oz + z;
        if (gi < 0 || gj < 0 || gk < 0 || gi >= d.nx || gj >= d.ny || gk >= d.nz) continue;
        const int c = x * SI + y * SJ + z;
        const float fc[6] = {FX[c], FX[c + SI], FY[c], FY[c + SJ], FZ[c], FZ[c + 1]};
        float v;
        if (ZI) {
            const float dg = fc[0] + fc[1] + fc[2] + fc[3] + fc[4] + fc[5];
            v = dg > 0.0f ? rb_first(B[c], dg, omega) : 0.0f;
        } else {
            const float pc = P[c], bc = B[c];
            const float pn[6] = {P[c - SI], P[c + SI], P[c - SJ], P[c + SJ], P[c - 1], P[c + 1]};
            float dg, Ap;
            stencil_vals(pc, fc, pn, dg, Ap);
            // (formed unconditionally: every read is issued at once)
            const float vr = rb_relax(pc, bc, Ap, dg, omega);
            v = dg > 0.0f ? vr : pc;
        }
        v = round_st<ST>(v);
        P[c] = v;
        // the ghosts (only the cell itself reads them; a low ghost of a cell
        // on the box's first layer lies outside the box, and that cell is
        // not updated again)
        if (gi == 0 && (M > 0 || x > 0)) P[c - SI] = v;
        if (gi == d.nx - 1) P[c + SI] = v;
        if (gj == 0 && (M > 0 || y > 0)) P[c - SJ] = v;
        if (gj == d.ny - 1) P[c + SJ] = v;
        if (gk == 0 && (M > 0 || z > 0)) P[c - 1] = v;
        if (gk == d.nz - 1) P[c + 1] = v;
    }
}

// half-sweeps h .. 2*NSW-1 at margins M0 + h, colours c1, 1 - c1, ...
template <typename ST, typename BX, int M0, int h, bool ZFIRST>
__device__ inline void lv_sweeps(float* P, const float* B, const float* FX, const float* FY, const float* FZ,
                                 const Dims& d, int ox, int oy, int oz, int c1, float omega) {
    if constexpr (h < 2 * (BX::H / 2)) {
        lv_half_sweep<ST, BX, M0 + h, ZFIRST && h == 0>(P, B, FX, FY, FZ, d, ox, oy, oz, (c1 + h) & 1, omega);
        __syncthreads();
        lv_sweeps<ST, BX, M0, h + 1, ZFIRST>(P, B, FX, FY, FZ, d, ox, oy, oz, c1, omega);
    }
}

// PRE: pin is unused (the level starts from zero), pout gets the smoothed
// iterate and bc the restricted residual (k_resid_restrict's order).
// !PRE: pout gets the post-smoothed iterate of pin + P pcoarse.
template <typename ST, bool PRE, int TI, int TJ, int TK, int NSW, typename CT = float>
__global__ __launch_bounds__(1024) void k_level(Dims d, Dims dc, const ST* __restrict__ bx,
                                                const ST* __restrict__ by, const ST* __restrict__ bz,
                                                const float* __restrict__ b, const ST* __restrict__ pin,
                                                const CT* __restrict__ pcoarse, ST* __restrict__ pout,
                                                float* __restrict__ bc, float omega,
                                                const int* __restrict__ done) {
    using BX = LvBox<TI, TJ, TK, NSW>;
    __shared__ float P[BX::N], B[BX::N], FX[BX::N], FY[BX::N], FZ[BX::N];
    const int tid = threadIdx.x;
    const int ntk = (d.nz + TK - 1) / TK, ntj = (d.ny + TJ - 1) / TJ;
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = (tile / (ntk * ntj)) * TI, j0 = ((tile / ntk) % ntj) * TJ, k0 = (tile % ntk) * TK;
    const int ox = i0 - BX::H, oy = j0 - BX::H, oz = k0 - BX::H;

    // staging: every load of the box issued before the first wait, all from
    // coordinates clamped into their array (no branches), the entries
    // outside it zeroed afterwards
    // (array indices by 32-bit 24x24-bit multiply-adds: the level's face
    // arrays hold fewer than 2^24 entries.  Left to itself the compiler
    // forms a*n + b as a 64-bit multiply-add whose unused upper addend half
    // may be the register of a load in flight, and each address then waits
    // for the previous load.)
    auto mad24 = [](unsigned a, unsigned n, unsigned b) {
        unsigned r;
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(n), "v"(b));
        return r;
    };
    auto ix = [&](unsigned a, unsigned n1, unsigned b, unsigned n2, unsigned c) {
        return mad24(mad24(a, n1, b), n2, c);
    };
    ST lf[BX::NPT][3], lp[BX::NPT];
    CT lc[BX::NPT];
    float lb[BX::NPT];
    unsigned m[BX::NPT];
#pragma unroll
    for (int q = 0; q < BX::NPT; ++q) {
        const int e = min(tid + q * BX::NT, BX::N - 1);
        const int z = e % BX::SJ, y = (e / BX::SJ) % (BX::BJ + 1), x = e / BX::SI;
        const int gi = ox + x, gj = oy + y, gk = oz + z;
        const int ci = min(max(gi, 0), d.nx - 1), cj = min(max(gj, 0), d.ny - 1), ck = min(max(gk, 0), d.nz - 1);
        const bool ii = gi == ci, jj = gj == cj, kk = gk == ck;
        // bit 0-2: the face of axis 0-2 exists, bit 3: the cell is in the grid
        m[q] = (tid + q * BX::NT < BX::N)
                   ? ((gi >= 0 && gi <= d.nx && jj && kk) | (ii && gj >= 0 && gj <= d.ny && kk) << 1 |
                      (ii && jj && gk >= 0 && gk <= d.nz) << 2 | (ii && jj && kk) << 3)
                   : 0u;
        const int fi = min(max(gi, 0), d.nx), fj = min(max(gj, 0), d.ny), fk = min(max(gk, 0), d.nz);
        lf[q][0] = bx[ix(fi, d.ny, cj, d.nz, ck)];
        lf[q][1] = by[ix(ci, d.ny + 1, fj, d.nz, ck)];
        lf[q][2] = bz[ix(ci, d.ny, cj, d.nz + 1, fk)];
        const unsigned cc = ix(ci, d.ny, cj, d.nz, ck);
        lb[q] = b[cc];
        if constexpr (!PRE) {
            lp[q] = pin[cc];
            lc[q] = pcoarse[ix(ci >> 1, dc.ny, cj >> 1, dc.nz, ck >> 1)];
        }
    }
    const bool skip = done && *done;
#pragma unroll
    for (int q = 0; q < BX::NPT; ++q) {
        const int e = tid + q * BX::NT;
        if (e < BX::N) {
            const unsigned mq = m[q];
            FX[e] = (mq & 1u) ? to_f32(lf[q][0]) : 0.0f;
            FY[e] = (mq & 2u) ? to_f32(lf[q][1]) : 0.0f;
            FZ[e] = (mq & 4u) ? to_f32(lf[q][2]) : 0.0f;
            B[e] = (mq & 8u) ? lb[q] : 0.0f;
            // (every entry from clamped coordinates: the ghosts hold their
            // cell's value)
            P[e] = PRE ? 0.0f : round_st<ST>(to_f32(lp[q]) + to_f32(lc[q]));
        }
    }
    __syncthreads();
    if (skip) return;  // a finished solve's queued iteration: no sweeps (uniform)
    // PRE: the first half-sweep from zero runs on the whole box (it reads no
    // neighbour), the last at margin H - 1; !PRE: margins 1 .. H
    lv_sweeps<ST, BX, PRE ? 0 : 1, 0, PRE>(P, B, FX, FY, FZ, d, ox, oy, oz, PRE ? 0 : 1, omega);

    // the tile goes out; PRE: each cell's residual (k_resid_restrict's
    // stencil with bscale 1) into LDS, then the sums of 2x2x2 children
    constexpr int TC = TI * TJ * TK;
    __shared__ float Rs[PRE ? TC : 1];
    for (int e = tid; e < TC; e += BX::NT) {
        const int kl = e % TK, jl = (e / TK) % TJ, il = e / (TK * TJ);
        const int gi = i0 + il, gj = j0 + jl, gk = k0 + kl;
        if (gi >= d.nx || gj >= d.ny || gk >= d.nz) continue;
        const int c = (il + BX::H) * BX::SI + (jl + BX::H) * BX::SJ + (kl + BX::H);
        const float pc = P[c];
        if (!skip) pout[cell_index(d, gi, gj, gk)] = from_f32<ST>(pc);
        if constexpr (PRE) {
            const float fc[6] = {FX[c], FX[c + BX::SI], FY[c], FY[c + BX::SJ], FZ[c], FZ[c + 1]};
            const float pn[6] = {P[c - BX::SI], P[c + BX::SI], P[c - BX::SJ], P[c + BX::SJ], P[c - 1], P[c + 1]};
            float dg, Ap;
            stencil_vals(pc, fc, pn, dg, Ap);
            Rs[e] = B[c] - Ap;
        }
    }
    if constexpr (PRE) {
        __syncthreads();
        constexpr int CJ = TJ / 2, CK = TK / 2;
        for (int e = tid; e < TC / 8; e += BX::NT) {
            const int K = e % CK, J = (e / CK) % CJ, I = e / (CK * CJ);
            const int gI = i0 / 2 + I, gJ = j0 / 2 + J, gK = k0 / 2 + K;
            if (gI >= dc.nx || gJ >= dc.ny || gK >= dc.nz) continue;
            float acc = 0.0f;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                    for (int c = 0; c < 2; ++c) acc += Rs[((2 * I + a) * TJ + 2 * J + bb) * TK + 2 * K + c];
            if (!skip) bc[cell_index(dc, gI, gJ, gK)] = acc;
        }
    }
}

// ================================================================ reductions
// Reductions are fence-free: a kernel writes one partial per block, and a
// one-block combine kernel (k_comb) forms the scalar.  The kernel boundary
// orders the two, so no device-scope fence (an L2 write-back per block on
// this multi-XCD part) is needed, and the fixed combine order keeps every
// result bitwise reproducible.
template <typename T>
__global__ __launch_bounds__(RT) void k_part_sum(const T* __restrict__ x, int64_t n, double* part) {
    __shared__ double sh[16];
    double v[1] = {0.0};
    for (int64_t c = (int64_t)blockIdx.x * RT + threadIdx.x; c < n; c += (int64_t)gridDim.x * RT)
        v[0] += (double)to_f32(x[c]);
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// y = x + shift (FP32 shift of the FP64 mean, as _remove_mean's shift_kernel)
template <typename TI, typename TO>
__global__ __launch_bounds__(RT) void k_shift(const TI* x, TO* y, int64_t n, const double* mean) {
    const float c = (float)(-*mean);
    for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT)
        y[i] = from_f32<TO>(to_f32(x[i]) + c);
}

template <typename T>
__global__ __launch_bounds__(RT) void k_zero(T* x, int64_t n, const int* __restrict__ done) {
    if (done && *done) return;
    for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT)
        x[i] = from_f32<T>(0.0f);
}

// ================================================================ small levels in LDS
struct SmallArgs {
    int s, L, pre, post, coarsest;
    float omega;
    Dims d[PF_MAX_LEVELS];
    int off[PF_MAX_LEVELS][5];  // LDS float offsets of bx, by, bz, p, b
    int scratch;                // LDS float offset of the residual scratch (n(s) floats)
    // global -> LDS copies at entry: segment q copies seg_len[q] floats from
    // seg_src[q] to LDS offset seg_dst[q] (the top level's faces and b, then
    // the coarser levels' faces and the coarsest level's matrix)
    int nseg;
    int nseg_first;  // segments [0, nseg_first): the top level's; the rest land during its descent
    const float* seg_src[3 * PF_MAX_LEVELS + 2];
    int seg_dst[3 * PF_MAX_LEVELS + 2];
    int seg_len[3 * PF_MAX_LEVELS + 2];
    void* p_top;
    const int* done;
    const double* cop;  // non-NULL: the coarsest level's smoother as a matrix
    int cop_lds;        // >= 0: LDS float offset of a staged copy of that matrix
};

constexpr int kSmallThreads = 1024;
constexpr int kSlots = 2;  // cells of one colour per thread and level

// Size of the per-colour enumeration of a level (pairs along k).  A level
// fits the single-workgroup V-cycle if it is at most kSmallThreads*kSlots;
// levels of at most 64 are smoothed by one wave.
__host__ __device__ inline int colour_space(const Dims& d) { return d.nx * d.ny * ((d.nz + 1) >> 1); }

// floor(t / d) for 0 <= t < 2^22 and 1 <= d, with inv = 1.0f / d: t + 1/2
// stays at least 1/(2d) away from a multiple of d, far more than the float
// rounding of the product.
__device__ inline int fdiv(int t, float inv) { return (int)(((float)t + 0.5f) * inv); }

struct LdsLevel {
    Dims d;
    float* bx;
    float* by;
    float* bz;
    float* p;
    float* b;
};

__device__ inline LdsLevel lds_level(const SmallArgs& a, float* lds, int l) {
    LdsLevel v;
    v.d = a.d[l];
    v.bx = lds + a.off[l][0];
    v.by = lds + a.off[l][1];
    v.bz = lds + a.off[l][2];
    v.p = lds + a.off[l][3];
    v.b = lds + a.off[l][4];
    return v;
}

// One cell that a thread updates on one level, decoded once per level visit:
// its index, the indices of its six neighbours with the reference's clamp at
// the walls (a missing neighbour is the cell itself), its y / z low faces
// (the x low face shares the cell index), its parent on the next coarser
// level, its right-hand side and omega / diag (0 where diag <= 0: no update).
struct Slot {
    int c, n[6], par;
    float f[6];  // x-low, x-high, y-low, y-high, z-low, z-high faces
    float b, w;
};

__device__ inline void make_slot(const LdsLevel& v, int colour, int t, float omega, const Dims* dc, Slot& S) {
    const Dims& d = v.d;
    const int nkp = (d.nz + 1) >> 1;
    S.c = -1;
    if (t >= d.nx * d.ny * nkp) return;
    const int row = fdiv(t, 1.0f / (float)nkp), kp = t - row * nkp;
    const int i = fdiv(row, 1.0f / (float)d.ny), j = row - i * d.ny;
    const int k = 2 * kp + ((i + j + colour) & 1);
    if (k >= d.nz) return;
    const int SI = d.ny * d.nz, SJ = d.nz;
    const int c = (i * d.ny + j) * d.nz + k;
    S.c = c;
    S.n[0] = i > 0 ? c - SI : c;
    S.n[1] = i < d.nx - 1 ? c + SI : c;
    S.n[2] = j > 0 ? c - SJ : c;
    S.n[3] = j < d.ny - 1 ? c + SJ : c;
    S.n[4] = k > 0 ? c - 1 : c;
    S.n[5] = k < d.nz - 1 ? c + 1 : c;
    const int fy = c + i * d.nz;        // face_index<1>(d, i, j, k)
    const int fz = c + (i * d.ny + j);  // face_index<2>(d, i, j, k)
    S.par = dc ? cell_index(*dc, i >> 1, j >> 1, k >> 1) : 0;
    S.f[0] = v.bx[c];
    S.f[1] = v.bx[c + SI];
    S.f[2] = v.by[fy];
    S.f[3] = v.by[fy + SJ];
    S.f[4] = v.bz[fz];
    S.f[5] = v.bz[fz + 1];
    const float dg = S.f[0] + S.f[1] + S.f[2] + S.f[3] + S.f[4] + S.f[5];
    S.b = v.b[c];
    S.w = dg > 0.0f ? omega / dg : 0.0f;
}

// A p at a slot's cell in the difference form (stencil_vals' operations)
__device__ inline float slot_Ap(const LdsLevel& v, const Slot& S, float pc) {
#pragma clang fp contract(off)
    float s = S.f[0] * (pc - v.p[S.n[0]]);
    s = fmaf(S.f[1], pc - v.p[S.n[1]], s);
    s = fmaf(S.f[2], pc - v.p[S.n[2]], s);
    s = fmaf(S.f[3], pc - v.p[S.n[3]], s);
    s = fmaf(S.f[4], pc - v.p[S.n[4]], s);
    s = fmaf(S.f[5], pc - v.p[S.n[5]], s);
    return s;
}

// One half-sweep over a thread's slots: every read first, then every write
// (cells of one colour never neighbour each other).
__device__ inline void slots_sweep(const LdsLevel& v, const Slot (&S)[kSlots]) {
    float nv[kSlots];
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        nv[q] = 0.0f;
        if (S[q].c < 0) continue;
        const float pc = v.p[S[q].c];
        nv[q] = pc + (S[q].b - slot_Ap(v, S[q], pc)) * S[q].w;
    }
#pragma unroll
    for (int q = 0; q < kSlots; ++q)
        if (S[q].c >= 0) v.p[S[q].c] = nv[q];
}

__device__ inline void wave_sync() {
    // LDS operations of one wave execute in issue order; this only keeps the
    // compiler from moving LDS accesses across the colour boundary
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup barrier over LDS only: unlike __syncthreads it does not wait
// for loads in flight, so the coarser levels' staging overlaps the top
// level's descent.
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The slots of both colours of one level for this thread (colour c0 first).
struct LevelSlots {
    Slot s0[kSlots], s1[kSlots];
    bool wave;  // smoothed by wave 0 alone
};

__device__ inline void make_level_slots(const LdsLevel& v, int c0, float omega, const Dims* dc, LevelSlots& L) {
    L.wave = colour_space(v.d) <= 64;
    const int nth = L.wave ? 64 : kSmallThreads;
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        const int t = threadIdx.x + q * nth;
        const bool mine = !L.wave || threadIdx.x < 64;
        L.s0[q].c = L.s1[q].c = -1;
        if (mine) {
            make_slot(v, c0, t, omega, dc, L.s0[q]);
            make_slot(v, 1 - c0, t, omega, dc, L.s1[q]);
        }
    }
}

// `sweeps` red-black sweeps (colour c0 first); ends with a workgroup barrier
__device__ inline void small_smooth(const LdsLevel& v, const LevelSlots& L, int sweeps) {
    if (L.wave) {
        if (threadIdx.x < 64)
            for (int sw = 0; sw < sweeps; ++sw) {
                slots_sweep(v, L.s0);
                wave_sync();
                slots_sweep(v, L.s1);
                wave_sync();
            }
    } else {
        for (int sw = 0; sw < sweeps; ++sw) {
            slots_sweep(v, L.s0);
            lds_barrier();
            slots_sweep(v, L.s1);
            lds_barrier();
        }
    }
    lds_barrier();
}

// The whole V-cycle below `s` in one workgroup, every level in LDS.
template <typename OT>
__global__ __launch_bounds__(kSmallThreads) void k_small_vcycle(SmallArgs a) {
    extern __shared__ float lds[];
    const int tid = threadIdx.x, nt = kSmallThreads;
    const int dn = a.done ? *a.done : 0;
    // segments straight from memory into LDS (lane l of a wave writes its
    // dword to the wave's LDS base + 4 l)
    auto stage = [&](int q0, int q1) {
        for (int q = q0; q < q1; ++q) {
            const float* src = a.seg_src[q];
            const int len = a.seg_len[q];
            float* dst = lds + a.seg_dst[q];
            for (int e0 = 0; e0 < len; e0 += nt)
                if (e0 + tid < len)
                    __builtin_amdgcn_global_load_lds(
                        (const void*)(src + e0 + tid), (__attribute__((address_space(3))) void*)(dst + e0 + (tid & ~63)),
                        4, 0, 0);
        }
    };
    // the top level's faces and b first; the coarser levels' faces and the
    // coarsest matrix are issued after that wait and land while the top
    // level is smoothed (its barriers are LDS-only), waited for before
    // their first use
    stage(0, a.nseg_first);
    __builtin_amdgcn_s_waitcnt(0);
    if (dn) return;  // uniform: no workgroup barrier has been reached yet
    stage(a.nseg_first, a.nseg);
    bool staged = a.nseg_first >= a.nseg;
    auto wait_staged = [&]() {
        if (!staged) {
            __builtin_amdgcn_s_waitcnt(0);
            staged = true;
        }
    };
    lds_barrier();
    float* scr = lds + a.scratch;
    LevelSlots S;
    // descend: smooth from zero, restrict the residual (poisson.py:323-336)
    for (int l = a.s; l < a.L - 1; ++l) {
        if (l > a.s) {
            wait_staged();
            lds_barrier();
        }
        const LdsLevel v = lds_level(a, lds, l);
        const LdsLevel c = lds_level(a, lds, l + 1);
        const int n = v.d.nx * v.d.ny * v.d.nz;
        for (int e = tid; e < n; e += nt) v.p[e] = 0.0f;
        make_level_slots(v, 0, a.omega, &c.d, S);
        lds_barrier();
        small_smooth(v, S, a.pre);
        // residual of every fine cell into the scratch, then R = P^T
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            if (S.s0[q].c >= 0) {
                const float pc = v.p[S.s0[q].c];
                scr[S.s0[q].c] = S.s0[q].b - slot_Ap(v, S.s0[q], pc);
            }
            if (S.s1[q].c >= 0) {
                const float pc = v.p[S.s1[q].c];
                scr[S.s1[q].c] = S.s1[q].b - slot_Ap(v, S.s1[q], pc);
            }
        }
        lds_barrier();
        const int nc = c.d.nx * c.d.ny * c.d.nz;
        const float inz = 1.0f / (float)c.d.nz, iny = 1.0f / (float)c.d.ny;
        for (int t = tid; t < nc; t += nt) {
            const int row = fdiv(t, inz), K = t - row * c.d.nz;
            const int I = fdiv(row, iny), J = row - I * c.d.ny;
            float acc = 0.0f;
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int z = 0; z < 2; ++z) acc += scr[cell_index(v.d, 2 * I + x, 2 * J + y, 2 * K + z)];
            c.b[t] = acc;
        }
        lds_barrier();
    }
    wait_staged();
    lds_barrier();
    // coarsest: sweeps, then remove the mean (poisson.py:325-328)
    if (a.cop) {
        // the same map, precomputed: p = M b (column j of M contiguous).
        // Each output's sum is split over G groups of n threads (a
        // contiguous range of j each; a wave reads 64 consecutive entries
        // of one column), and the G partials are added in group order
        // (the FP64 partials go through the scratch, 8-byte aligned).
        const LdsLevel v = lds_level(a, lds, a.L - 1);
        const int n = v.d.nx * v.d.ny * v.d.nz;
        const int G = min(nt / n, 16), J = (n + G - 1) / G;
        double* part = reinterpret_cast<double*>(scr);
        // (one instance per memory kind, so the loads are LDS or global
        // loads, not generic ones)
        auto partial = [&](const auto* M) {
            if (tid < n * G) {
                const int t = tid % n, g = tid / n;
                const int j1 = min(n, (g + 1) * J);
                double acc = 0.0;
#pragma unroll 4
                for (int j = g * J; j < j1; ++j) acc += M[(int64_t)j * n + t] * (double)v.b[j];
                part[g * n + t] = acc;
            }
        };
        // the matrix staged in LDS with the faces, else read from memory
        if (a.cop_lds >= 0)
            partial(reinterpret_cast<const double*>(lds + a.cop_lds));
        else
            partial(a.cop);
        lds_barrier();
        for (int t = tid; t < n; t += nt) {
            double acc = part[t];
            for (int g = 1; g < G; ++g) acc += part[g * n + t];
            v.p[t] = (float)acc;
        }
        lds_barrier();
    } else {
        const LdsLevel v = lds_level(a, lds, a.L - 1);
        const int n = v.d.nx * v.d.ny * v.d.nz;
        for (int e = tid; e < n; e += nt) v.p[e] = 0.0f;
        make_level_slots(v, 0, a.omega, nullptr, S);
        __syncthreads();
        small_smooth(v, S, a.coarsest);
        __shared__ double sh[16];
        double sm[1] = {0.0};
        for (int e = tid; e < n; e += nt) sm[0] += (double)v.p[e];
        block_sum<1>(sm, sh);
        __shared__ float shift;
        if (tid == 0) shift = (float)(-(sm[0] / (double)n));
        __syncthreads();
        for (int e = tid; e < n; e += nt) v.p[e] = v.p[e] + shift;
        __syncthreads();
    }
    // ascend: prolong (poisson.py:178-182), smooth (colours (1,0))
    for (int l = a.L - 2; l >= a.s; --l) {
        const LdsLevel v = lds_level(a, lds, l);
        const LdsLevel c = lds_level(a, lds, l + 1);
        make_level_slots(v, 1, a.omega, &c.d, S);
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            if (S.s0[q].c >= 0) v.p[S.s0[q].c] += c.p[S.s0[q].par];
            if (S.s1[q].c >= 0) v.p[S.s1[q].c] += c.p[S.s1[q].par];
        }
        __syncthreads();
        small_smooth(v, S, a.post);
    }
    {
        const LdsLevel v = lds_level(a, lds, a.s);
        const int n = v.d.nx * v.d.ny * v.d.nz;
        OT* out = static_cast<OT*>(a.p_top);
        for (int e = tid; e < n; e += nt) out[e] = from_f32<OT>(v.p[e]);
    }
}

// Column j of the coarsest level's smoother map: `sweeps` red-black sweeps
// (colours (0,1)) from p = 0 with b = e_j, then mean removal, as the
// single-workgroup V-cycle runs them, in FP64.  One workgroup per column,
// one thread per cell (n <= kCoarseOpMax).
__global__ __launch_bounds__(kCoarseOpMax) void k_coarse_op(Dims d, const float* __restrict__ bx,
                                                            const float* __restrict__ by,
                                                            const float* __restrict__ bz, int sweeps, float omega,
                                                            double* __restrict__ op) {
    __shared__ double P[kCoarseOpMax];
    __shared__ double sh[16];
    __shared__ double shift;
    const int n = d.nx * d.ny * d.nz, j = blockIdx.x, t = threadIdx.x;
    const bool own = t < n;
    double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int nb[6] = {0, 0, 0, 0, 0, 0};
    int colour = 0;
    double dg = 0.0;
    if (own) {
        const int k = t % d.nz, jj = (t / d.nz) % d.ny, i = t / (d.nz * d.ny);
        f[0] = bx[face_index<0>(d, i, jj, k)];
        f[1] = bx[face_index<0>(d, i + 1, jj, k)];
        f[2] = by[face_index<1>(d, i, jj, k)];
        f[3] = by[face_index<1>(d, i, jj + 1, k)];
        f[4] = bz[face_index<2>(d, i, jj, k)];
        f[5] = bz[face_index<2>(d, i, jj, k + 1)];
        nb[0] = cell_index(d, max(i - 1, 0), jj, k);
        nb[1] = cell_index(d, min(i + 1, d.nx - 1), jj, k);
        nb[2] = cell_index(d, i, max(jj - 1, 0), k);
        nb[3] = cell_index(d, i, min(jj + 1, d.ny - 1), k);
        nb[4] = cell_index(d, i, jj, max(k - 1, 0));
        nb[5] = cell_index(d, i, jj, min(k + 1, d.nz - 1));
        colour = (i + jj + k) & 1;
        dg = f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
        P[t] = 0.0;
    }
    const double b = t == j ? 1.0 : 0.0, w = (double)omega;
    __syncthreads();
    for (int sw = 0; sw < sweeps; ++sw) {
        for (int c = 0; c < 2; ++c) {
            const bool upd = own && colour == c && dg > 0.0;
            double v = 0.0;
            if (upd) {
                const double pc = P[t];
                double ap = 0.0;
#pragma unroll
                for (int q = 0; q < 6; ++q) ap += f[q] * (pc - P[nb[q]]);
                v = pc + w * (b - ap) / dg;
            }
            __syncthreads();
            if (upd) P[t] = v;
            __syncthreads();
        }
    }
    double sm[1] = {own ? P[t] : 0.0};
    block_sum<1>(sm, sh);
    if (t == 0) shift = -(sm[0] / (double)n);
    __syncthreads();
    if (own) op[(int64_t)j * n + t] = P[t] + shift;
}

static bool use_coarse_op(const pf_mg& mg) {
    const pf_mg_level& c = mg.lev[mg.nlev - 1];
    return mg.small_from < mg.nlev && (int64_t)c.nx * c.ny * c.nz <= kCoarseOpMax;
}

// ================================================================ FPCG kernels
template <typename ZT>
struct CGK {
    Dims d;
    int64_t n;
    const float* bx;
    const float* by;
    const float* bz;
    const float* b;
    float* rhs;
    float* x;
    float* r;
    float* q;            // search direction written by this iteration
    const float* q_old;  // the previous one (the two alternate between iterations)
    float* Ap;
    const ZT* z;
    const float* zscale;  // NULL: z as stored; else z times *zscale
    CGState* st;
    double tol;
    int slot;      // parity of the iteration (selects q / q_old and rz)
    int np_apply;  // blocks of k_cg_apply (its <q, Aq> partials)
    int np_beta;   // blocks of k_cg_beta (its dot-product partials)
};

template <typename ZT>
__device__ inline float zval(const CGK<ZT>& a, float zs, int64_t i) {
    return to_f32(a.z[i]) * zs;
}

// r = b - mean(b); x = 0; q = 0; partial max|r| (poisson.py:363-372)
template <typename ZT>
__global__ __launch_bounds__(RT) void k_cg_init(CGK<ZT> a) {
    __shared__ float sh[4];
    const float c = (float)(-a.st->mu);
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * RT) {
        const float v = a.b[i] + c;
        a.r[i] = v;
        a.x[i] = 0.0f;
        a.q[i] = 0.0f;  // the first iteration's previous direction (beta = 0)
        m = fmaxf(m, fabsf(v));
    }
    m = block_max(m, sh);
    if (threadIdx.x == 0) a.st->fpart[blockIdx.x] = m;
}

// partials of sum(z), sum(r), <r, z> after the first V-cycle
template <typename ZT>
__global__ __launch_bounds__(RT) void k_cg_first(CGK<ZT> a) {
    __shared__ double sh[3 * 16];
    if (a.st->done) return;
    const float zs = a.zscale ? *a.zscale : 1.0f;
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * RT) {
        const double zz = (double)zval(a, zs, i), rr = (double)a.r[i];
        v[0] += zz;
        v[1] += rr;
        v[2] += rr * zz;
    }
    block_sum<3>(v, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) a.st->part[q * MAXB + blockIdx.x] = v[q];
}

// q = (z - mu) + beta q_old (the search direction, poisson.py:406; beta = 0
// on the first iteration), evaluated at the cell and its six neighbours;
// then Ap = A q and the partial <q, Ap> (poisson.py:385-386).  Writes q and Ap
// of the cell; q_old is a different buffer, so neighbours read old values.
// From the second iteration on it first combines k_cg_beta's partials into
// the mean of z and the Polak-Ribiere beta (poisson.py:400-406) itself:
// every block sums them in the same fixed order, so all blocks hold the same
// scalars, and block 0 records them (no separate combine launch).
template <typename ZT>
__global__ __launch_bounds__(WT) void k_cg_apply(CGK<ZT> a) {
    __shared__ double sh[5 * 16];
    __shared__ float s_cm, s_be;
    CGState* st = a.st;
    const Dims d = a.d;
    // 32-bit indices (the level's face arrays hold fewer than 2^31
    // entries): the thread's first cell and the grid stride are decoded into
    // (i, j, k) once, and the loop adds the stride with carries; the faces
    // and the clamped neighbours are offsets from the cell index
    // (face_index and cell_index written out).  64-bit division per cell
    // made this loop several times longer.
    const int n = (int)a.n, sjk = d.ny * d.nz;
    const int c0 = blockIdx.x * WT + threadIdx.x, S = gridDim.x * WT;
    int k = c0 % d.nz, j = (c0 / d.nz) % d.ny, i = c0 / sjk;
    const int sk = S % d.nz, sj = (S / d.nz) % d.ny, si = S / sjk;
    // a cell's inputs: faces, then z (as FP32, unscaled) and q_old at the
    // cell and its six clamped neighbours (-x, +x, -y, +y, -z, +z)
    float f[6], zr[7], qo[7];
    auto load_cell = [&](int c) {
        const int fy = c + i * d.nz, fz = c + i * d.ny + j;
        f[0] = a.bx[c];
        f[1] = a.bx[c + sjk];
        f[2] = a.by[fy];
        f[3] = a.by[fy + d.nz];
        f[4] = a.bz[fz];
        f[5] = a.bz[fz + 1];
        const int nb[7] = {c, i > 0 ? c - sjk : c, i < d.nx - 1 ? c + sjk : c, j > 0 ? c - d.nz : c,
                           j < d.ny - 1 ? c + d.nz : c, k > 0 ? c - 1 : c, k < d.nz - 1 ? c + 1 : c};
#pragma unroll
        for (int t = 0; t < 7; ++t) {
            zr[t] = to_f32(a.z[nb[t]]);
            qo[t] = a.q_old[nb[t]];
        }
    };
    // the first cell's loads are issued before the flag test and the
    // combine below, so their round trip overlaps that work (a finished
    // solve's queued launch returns with them in flight)
    if (c0 < n) load_cell(c0);
    const float zs = a.zscale ? *a.zscale : 1.0f;
    if (st->done) return;
    if (st->beta_parts) {
        // z' = z - mean(z); rz_new = <r, z'>; num = rz_new - <r_prev, z'> =
        // -alpha <Ap, z'> because r_prev = r + alpha Ap
        double t[5];
        combine_sums<5>(st->part, a.np_beta, t, sh);
        if (threadIdx.x == 0) {
            const double mu = t[0] / (double)a.n;
            const double rz_new = t[3] - mu * t[1];
            const double apz = t[4] - mu * t[2];
            const double num = -st->alpha * apz;
            const double rz = st->rz[a.slot ^ 1];
            const double beta = (rz != 0.0) ? fmax(num / rz, 0.0) : 0.0;
            s_cm = (float)(-mu);
            s_be = (float)beta;
            if (blockIdx.x == 0) {
                st->mu = mu;
                st->beta = beta;
                st->rz[a.slot] = rz_new;
            }
        }
    } else if (threadIdx.x == 0) {  // the first iteration: k_comb(C_FIRST) set them
        s_cm = (float)(-st->mu);
        s_be = (float)st->beta;
    }
    __syncthreads();
    const float cm = s_cm;
    const float be = s_be;
    double v[1] = {0.0};
    // q at input point t: (z - mu) + beta q_old
    auto qv = [&](int t) {
        const float zz = zr[t] * zs + cm;
        return zz + be * qo[t];
    };
    for (int c = c0; c < n; c += S) {
        if (c != c0) load_cell(c);
        const float pc = qv(0);
        const float pn[6] = {qv(1), qv(2), qv(3), qv(4), qv(5), qv(6)};
        float dg, Ap;
        stencil_vals(pc, f, pn, dg, Ap);
        a.q[c] = pc;
        a.Ap[c] = Ap;
        v[0] += (double)pc * (double)Ap;
        k += sk;
        if (k >= d.nz) {
            k -= d.nz;
            ++j;
        }
        j += sj;
        if (j >= d.ny) {
            j -= d.ny;
            ++i;
        }
        i += si;
    }
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) st->part[5 * MAXB + blockIdx.x] = v[0];
}

// x += alpha q; r -= alpha Ap; partial max|r| (poisson.py:390-394), after
// combining k_cg_apply's partials into <q, Aq> and alpha = rz / <q, Aq> in
// every block (fixed order, so all blocks agree); <q, Aq> <= 0 ends the
// solve before any update (poisson.py:386-389).  Block 0 records the
// scalars and the iteration count.
template <typename ZT>
__global__ __launch_bounds__(RT) void k_cg_update(CGK<ZT> a) {
    __shared__ float sh[4];
    __shared__ double shd[16];
    __shared__ float s_al;
    __shared__ int s_ok;
    CGState* st = a.st;
    // the first element's inputs are loaded before the flag test and the
    // combine, so their round trip overlaps that work
    const int64_t i0 = (int64_t)blockIdx.x * RT + threadIdx.x, stride = (int64_t)gridDim.x * RT;
    float x0 = 0.f, q0 = 0.f, r0 = 0.f, ap0 = 0.f;
    if (i0 < a.n) {
        x0 = a.x[i0];
        q0 = a.q[i0];
        r0 = a.r[i0];
        ap0 = a.Ap[i0];
    }
    if (st->done) return;
    double t[1];
    combine_sums<1>(st->part + 5 * MAXB, a.np_apply, t, shd);
    if (threadIdx.x == 0) {
        const double qAq = t[0];
        const bool ok = qAq > 0.0;
        const double alpha = ok ? st->rz[a.slot] / qAq : 0.0;
        s_ok = ok ? 1 : 0;
        s_al = (float)alpha;
        if (blockIdx.x == 0) {
            st->qAq = qAq;
            st->it += 1;
            if (ok)
                st->alpha = alpha;
            else
                st->done = 1;
        }
    }
    __syncthreads();
    if (!s_ok) return;
    const float al = s_al, nal = -s_al;
    float m = 0.0f;
    for (int64_t i = i0; i < a.n; i += stride) {
        if (i != i0) {
            x0 = a.x[i];
            q0 = a.q[i];
            r0 = a.r[i];
            ap0 = a.Ap[i];
        }
        a.x[i] = x0 + al * q0;
        const float rn = r0 + nal * ap0;
        a.r[i] = rn;
        m = fmaxf(m, fabsf(rn));
    }
    m = block_max(m, sh);
    if (threadIdx.x == 0) a.st->fpart[blockIdx.x] = m;
}

// partials of sum(z), sum(r), sum(Ap), <r,z>, <Ap,z> after a V-cycle
template <typename ZT>
__global__ __launch_bounds__(WT) void k_cg_beta(CGK<ZT> a) {
    __shared__ double sh[5 * 16];
    if (a.st->done) return;
    const float zs = a.zscale ? *a.zscale : 1.0f;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * WT + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * WT) {
        const double zz = (double)zval(a, zs, i), rr = (double)a.r[i], ap = (double)a.Ap[i];
        v[0] += zz;
        v[1] += rr;
        v[2] += ap;
        v[3] += rr * zz;
        v[4] += ap * zz;
    }
    block_sum<5>(v, sh);
    if (threadIdx.x == 0) {
        for (int q = 0; q < 5; ++q) a.st->part[q * MAXB + blockIdx.x] = v[q];
        if (blockIdx.x == 0) a.st->beta_parts = 1;  // read by the next k_cg_apply
    }
}

// One-block combine of the previous kernel's partials, then the scalar step.
enum CombOp { C_MEAN = 0, C_MEAN_RESET, C_INIT, C_FIRST, C_UPDATE };


__global__ __launch_bounds__(RT) void k_comb(CGState* st, int op, int nparts, int64_t n,
                                             double* out, double tol, int* host_flag = nullptr) {
    __shared__ double sh[5 * 16];
    __shared__ float shf[4];
    if (op == C_MEAN || op == C_MEAN_RESET) {
        double t[1];
        combine_sums<1>(st->part, nparts, t, sh);
        if (threadIdx.x == 0) {
            *out = t[0] / (double)n;
            if (op == C_MEAN_RESET) {
                st->done = 0;
                st->it = 0;
                st->seq = 0;
                st->beta_parts = 0;
                st->rel = 1.0;
            }
        }
        return;
    }
    if (op == C_INIT) {
        // b_norm = max|rhs|; b == 0 returns x = 0 (poisson.py:365-368)
        const float mm = combine_max(st->fpart, nparts, shf);
        if (threadIdx.x == 0) {
            st->b_norm = (double)mm;
            st->tol = tol;
            set_vscale(st, mm);
            if (mm == 0.0f) {
                st->done = 2;
                st->rel = 0.0;
            }
        }
        return;
    }
    if (op == C_UPDATE) {
        // rel = max|r| / max|b| (poisson.py:394-398)
        const float mm = st->done ? 0.0f : combine_max(st->fpart, nparts, shf);
        if (threadIdx.x == 0) residual_test_record(st, mm, host_flag);
        return;
    }
    if (st->done) return;
    if (op == C_FIRST) {
        // mu = mean(z); rz = <r, z - mu>
        double t[3];
        combine_sums<3>(st->part, nparts, t, sh);
        if (threadIdx.x == 0) {
            // iteration 1 has parity 1: its k_cg_update reads rz[1]
            const double mu = t[0] / (double)n;
            st->mu = mu;
            st->rz[1] = t[2] - mu * t[1];
            st->beta = 0.0;
            st->beta_parts = 0;
        }
    }
}

__global__ __launch_bounds__(RT) void k_apply_A(Dims d, const float* __restrict__ p,
                                                const float* __restrict__ bx,
                                                const float* __restrict__ by,
                                                const float* __restrict__ bz,
                                                const float* __restrict__ diag, float* out) {
    const int64_t c = (int64_t)blockIdx.x * RT + threadIdx.x;
    if (c >= (int64_t)d.nx * d.ny * d.nz) return;
    const int k = (int)(c % d.nz), j = (int)((c / d.nz) % d.ny), i = (int)(c / ((int64_t)d.nz * d.ny));
    float dg, sn, pc;
    stencil_s(d, bx, by, bz, p, i, j, k, dg, sn, pc);
    out[c] = (diag ? diag[c] : dg) * pc - sn;
}

// ================================================================ host side
static Dims ldims(const pf_mg_level& l) { return Dims{l.nx, l.ny, l.nz}; }
static int64_t ncell(const pf_mg_level& l) { return (int64_t)l.nx * l.ny * l.nz; }

// LDS layout of the single-workgroup V-cycle from level s down: the face
// arrays of every level back to back (one contiguous copy), then p and b of
// every level, then a residual scratch of level s's size.
static int64_t small_lds_floats(const pf_mg& mg, int s, int (*off)[5], int* scratch) {
    int64_t o = 0;
    for (int l = s; l < mg.nlev; ++l) {
        const Dims d = ldims(mg.lev[l]);
        if (off) off[l][0] = (int)o;
        o += face_count(d, 0);
        if (off) off[l][1] = (int)o;
        o += face_count(d, 1);
        if (off) off[l][2] = (int)o;
        o += face_count(d, 2);
    }
    for (int l = s; l < mg.nlev; ++l) {
        const int64_t nc = ncell(mg.lev[l]);
        if (off) off[l][3] = (int)o;
        o += nc;
        if (off) off[l][4] = (int)o;
        o += nc;
    }
    // the scratch: the top level's residual, and the FP64 partials of the
    // coarsest level's matrix product (8-byte aligned, at most kSmallThreads)
    o = (o + 1) & ~(int64_t)1;
    if (scratch) *scratch = (int)o;
    return o + std::max<int64_t>(ncell(mg.lev[s]), 2 * kSmallThreads);
}

constexpr int64_t kSmallMaxCells = 8192;
constexpr int64_t kSmallMaxBytes = 150 * 1024;

template <typename ST>
struct Lv {
    static const ST* fx(const pf_mg_level& l) { return reinterpret_cast<const ST*>(l.bx_h); }
    static const ST* fy(const pf_mg_level& l) { return reinterpret_cast<const ST*>(l.by_h); }
    static const ST* fz(const pf_mg_level& l) { return reinterpret_cast<const ST*>(l.bz_h); }
    static ST* p(const pf_mg_level& l) { return reinterpret_cast<ST*>(l.p_h); }
};
template <>
struct Lv<float> {
    static const float* fx(const pf_mg_level& l) { return l.bx; }
    static const float* fy(const pf_mg_level& l) { return l.by; }
    static const float* fz(const pf_mg_level& l) { return l.bz; }
    static float* p(const pf_mg_level& l) { return l.p; }
};

// The single-workgroup kernel needs more than the default 64 KB of dynamic
// LDS; raise the limit once per device (gfx950 has 160 KB per workgroup).
static int ensure_small_lds() {
    static std::atomic<uint64_t> done_mask{0};
    int dev = 0;
    PF_CHECK_HIP(hipGetDevice(&dev));
    const uint64_t bit = 1ull << (dev & 63);
    if (done_mask.load() & bit) return PF_OK;
    const int bytes = (int)kSmallMaxBytes;
    PF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small_vcycle<float>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    PF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small_vcycle<bf16s>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    PF_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small_vcycle<fp16s>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    done_mask.fetch_or(bit);
    return PF_OK;
}

// Fused full sweep on one level: tile shape by level size (big tiles cut
// the redundant halo work, small ones keep enough workgroups in flight).
// PROL (CT = the coarse iterate's type): the input is pin plus the
// prolongation of pcoarse (dims dc); the fine dims must be twice dc.
template <typename ST, bool ZINIT, bool PROL = false, typename CT = float>
static int launch_full_sweep(const Dims& d, const ST* fx, const ST* fy, const ST* fz, const ST* pin,
                             ST* pout, const float* b, int c1, float omega, const int* done,
                             const float* bscale, hipStream_t s, const CT* pcoarse = nullptr,
                             Dims dc = Dims{1, 1, 1}, UpdateFold uf = UpdateFold{nullptr, 0, nullptr},
                             DotFold df = DotFold{nullptr, nullptr, nullptr}) {
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    auto tiles = [&](int ti, int tj, int tk) {
        return (unsigned)(cdiv(d.nx, ti) * cdiv(d.ny, tj) * cdiv(d.nz, tk));
    };
    if constexpr (!ZINIT && !PROL) {
        if (df.Ap) {  // the plain sweep with the CG's inner products
            if (n >= (int64_t)1 << 20)
                k_rbgs_full<ST, false, 8, 8, 64, 512, false, float, true><<<tiles(8, 8, 64), 512, 0, s>>>(
                    d, fx, fy, fz, pin, pout, b, c1, omega, done, bscale, nullptr, dc, uf, df);
            else if (n >= (int64_t)1 << 17)
                k_rbgs_full<ST, false, 8, 8, 16, 256, false, float, true><<<tiles(8, 8, 16), 256, 0, s>>>(
                    d, fx, fy, fz, pin, pout, b, c1, omega, done, bscale, nullptr, dc, uf, df);
            else
                k_rbgs_full<ST, false, 4, 8, 16, 256, false, float, true><<<tiles(4, 8, 16), 256, 0, s>>>(
                    d, fx, fy, fz, pin, pout, b, c1, omega, done, bscale, nullptr, dc, uf, df);
            PF_CHECK_LAUNCH();
            return PF_OK;
        }
    }
    PF_REQUIRE(!df.Ap, "launch_full_sweep: inner products only on a plain sweep");
    if (n >= (int64_t)1 << 20)
        k_rbgs_full<ST, ZINIT, 8, 8, 64, 512, PROL, CT><<<tiles(8, 8, 64), 512, 0, s>>>(
            d, fx, fy, fz, pin, pout, b, c1, omega, done, bscale, pcoarse, dc, uf, df);
    else if (n >= (int64_t)1 << 17)
        k_rbgs_full<ST, ZINIT, 8, 8, 16, 256, PROL, CT><<<tiles(8, 8, 16), 256, 0, s>>>(
            d, fx, fy, fz, pin, pout, b, c1, omega, done, bscale, pcoarse, dc, uf, df);
    else
        k_rbgs_full<ST, ZINIT, 4, 8, 16, 256, PROL, CT><<<tiles(4, 8, 16), 256, 0, s>>>(
            d, fx, fy, fz, pin, pout, b, c1, omega, done, bscale, pcoarse, dc, uf, df);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

// Largest level whose pre and post phases run as one k_level launch each
// (32^3: 256 tiles of 4 x 4 x 8 cells, one workgroup per CU).
constexpr int64_t kLevelFusedMaxCells = 32768;

// One k_level launch: PRE (the pre-smoothing phase from zero and the
// restricted residual into bc) or the post phase (pin plus the
// prolongation of pcoarse, dims dc, into pout); nsw sweeps, 1..3.
template <typename ST, bool PRE, typename CT = float>
static int launch_level(const Dims& d, const Dims& dc, int nsw, const ST* fx, const ST* fy, const ST* fz,
                        const float* b, const ST* pin, const CT* pcoarse, ST* pout, float* bc, float omega,
                        const int* done, hipStream_t s) {
    constexpr int TI = 4, TJ = 4, TK = 8;
    PF_REQUIRE(d.nx == 2 * dc.nx && d.ny == 2 * dc.ny && d.nz == 2 * dc.nz, "launch_level: fine dims must be 2x coarse");
    // the kernel forms indices with 24-bit multiplies
    PF_REQUIRE((int64_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1) < ((int64_t)1 << 24), "launch_level: level too large");
    const unsigned grid = cdiv(d.nx, TI) * cdiv(d.ny, TJ) * cdiv(d.nz, TK);
    switch (nsw) {
    case 1:
        k_level<ST, PRE, TI, TJ, TK, 1, CT><<<grid, 1024, 0, s>>>(d, dc, fx, fy, fz, b, pin, pcoarse, pout, bc,
                                                                  omega, done);
        break;
    case 2:
        k_level<ST, PRE, TI, TJ, TK, 2, CT><<<grid, 1024, 0, s>>>(d, dc, fx, fy, fz, b, pin, pcoarse, pout, bc,
                                                                  omega, done);
        break;
    case 3:
        k_level<ST, PRE, TI, TJ, TK, 3, CT><<<grid, 1024, 0, s>>>(d, dc, fx, fy, fz, b, pin, pcoarse, pout, bc,
                                                                  omega, done);
        break;
    default:
        PF_REQUIRE(false, "launch_level: 1 to 3 sweeps");
    }
    PF_CHECK_LAUNCH();
    return PF_OK;
}

// workgroups of launch_full_sweep on a level of dims d (one per tile)
static int64_t full_sweep_blocks(const Dims& d) {
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    const int ti = n >= (int64_t)1 << 20 ? 8 : (n >= (int64_t)1 << 17 ? 8 : 4);
    const int tk = n >= (int64_t)1 << 20 ? 64 : 16;
    return cdiv(d.nx, ti) * cdiv(d.ny, 8) * cdiv(d.nz, tk);
}

// The two iterate buffers of an HBM level: the level's own (level 0: the
// output z0) and the fused smoother's second buffer.
template <typename ST>
static ST* second_buffer(const pf_mg& mg, int l) {
    if (sizeof(ST) == 4 && l == 0) return reinterpret_cast<ST*>(mg.lev[0].p);
    return reinterpret_cast<ST*>(mg.lev[l].r);
}

static bool fused_smoother(const pf_mg& mg, int l) {
    return !(mg.options & PF_MG_UNFUSED_SMOOTHER) && (mg.lev[l].nz % 2) == 0;
}

// level l (restricted to l + 1, below level 0) runs each smoothing phase
// as one k_level launch
static bool level_fused(const pf_mg& mg, int l, int G) {
    return l >= 1 && l < G && !(mg.options & (PF_MG_UNFUSED_SMOOTHER | PF_MG_SPLIT_LEVELS)) &&
           ncell(mg.lev[l]) <= kLevelFusedMaxCells && mg.pre >= 1 && mg.pre <= 3 && mg.post >= 1 &&
           mg.post <= 3;
}

// One V-cycle: z0 = M^{-1} r0 (no mean removal at level 0).
template <typename ST>
// dfold: the CG's inner products ride on the last level-0 post-sweep
// (*dot_done tells whether one took them).
static int vcycle(const pf_mg& mg, const float* r0, ST* z0, const int* done, const float* bscale,
                  hipStream_t s, const UpdateFold* uf = nullptr, const DotFold* dfold = nullptr,
                  bool* dot_done = nullptr) {
    const int L = mg.nlev;
    const int sm = mg.small_from;             // first LDS level (L if none)
    const int G = sm < L ? sm : L - 1;        // levels 0..G-1 run out of HBM with restriction
    auto P = [&](int l) -> ST* { return l == 0 ? z0 : Lv<ST>::p(mg.lev[l]); };
    auto B = [&](int l) -> const float* { return l == 0 ? r0 : mg.lev[l].b; };
    ST* cur[PF_MAX_LEVELS];  // buffer holding each HBM level's current iterate
    auto other = [&](int l, const ST* x) -> ST* { return x == P(l) ? second_buffer<ST>(mg, l) : P(l); };

    // `sweeps` red-black sweeps of level l (colour order (0,1), or (1,0) if
    // reverse), starting from zero if from_zero, else from cur[l]
    auto smooth = [&](int l, int sweeps, bool reverse, bool from_zero) -> int {
        const pf_mg_level& lv = mg.lev[l];
        const Dims d = ldims(lv);
        const ST *fx = Lv<ST>::fx(lv), *fy = Lv<ST>::fy(lv), *fz = Lv<ST>::fz(lv);
        const float* b = B(l);
        const float* bsc = l == 0 ? bscale : nullptr;
        if (from_zero) cur[l] = P(l);
        if (from_zero && sweeps == 0) {
            k_zero<ST><<<red_blocks(ncell(lv)), RT, 0, s>>>(cur[l], ncell(lv), done);
            PF_CHECK_LAUNCH();
        }
        if (fused_smoother(mg, l)) {
            for (int sw = 0; sw < sweeps; ++sw) {
                ST* out = other(l, cur[l]);
                const int c1 = reverse ? 1 : 0;
                // the V-cycle's first kernel takes the residual test with it
                const UpdateFold f = (uf && l == 0 && sw == 0) ? *uf : UpdateFold{nullptr, 0, nullptr};
                const bool dots = dfold && l == 0 && !from_zero && sw == sweeps - 1;
                const int rc = (from_zero && sw == 0)
                                   ? launch_full_sweep<ST, true>(d, fx, fy, fz, cur[l], out, b, c1, mg.omega, done,
                                                                 bsc, s, static_cast<const float*>(nullptr),
                                                                 Dims{1, 1, 1}, f)
                                   : launch_full_sweep<ST, false>(d, fx, fy, fz, cur[l], out, b, c1, mg.omega, done,
                                                                  bsc, s, static_cast<const float*>(nullptr),
                                                                  Dims{1, 1, 1}, UpdateFold{nullptr, 0, nullptr},
                                                                  dots ? *dfold : DotFold{nullptr, nullptr, nullptr});
                if (dots && dot_done) *dot_done = true;
                if (rc) return rc;
                cur[l] = out;
            }
            return PF_OK;
        }
        const int64_t nthr = (int64_t)d.nx * d.ny * ((d.nz + 1) / 2);
        const unsigned grid = cdiv(nthr, RT);
        for (int sw = 0; sw < sweeps; ++sw) {
            for (int h = 0; h < 2; ++h) {
                const int colour = reverse ? 1 - h : h;
                if (from_zero && sw == 0 && h == 0)
                    k_rbgs<ST, true><<<grid, RT, 0, s>>>(d, fx, fy, fz, cur[l], b, colour, mg.omega, done, bsc);
                else
                    k_rbgs<ST, false><<<grid, RT, 0, s>>>(d, fx, fy, fz, cur[l], b, colour, mg.omega, done, bsc);
                PF_CHECK_LAUNCH();
            }
        }
        return PF_OK;
    };

    for (int l = 0; l < G; ++l) {
        const pf_mg_level& lv = mg.lev[l];
        const pf_mg_level& cv = mg.lev[l + 1];
        if (level_fused(mg, l, G)) {
            cur[l] = P(l);
            const int rc = launch_level<ST, true>(ldims(lv), ldims(cv), mg.pre, Lv<ST>::fx(lv), Lv<ST>::fy(lv),
                                                  Lv<ST>::fz(lv), B(l), static_cast<const ST*>(nullptr),
                                                  static_cast<const float*>(nullptr), cur[l], mg.lev[l + 1].b,
                                                  mg.omega, done, s);
            if (rc) return rc;
            continue;
        }
        int rc = smooth(l, mg.pre, false, true);
        if (rc) return rc;
        k_resid_restrict<ST><<<cdiv(ncell(cv), RT), RT, 0, s>>>(ldims(lv), ldims(cv), Lv<ST>::fx(lv),
                                                              Lv<ST>::fy(lv), Lv<ST>::fz(lv), cur[l], B(l),
                                                              mg.lev[l + 1].b, done, l == 0 ? bscale : nullptr);
        PF_CHECK_LAUNCH();
    }
    if (sm < L) {
        int rc0 = ensure_small_lds();
        if (rc0) return rc0;
        SmallArgs a;
        a.s = sm;
        a.L = L;
        a.pre = mg.pre;
        a.post = mg.post;
        a.coarsest = mg.coarsest;
        a.omega = mg.omega;
        for (int l = sm; l < L; ++l) a.d[l] = ldims(mg.lev[l]);
        const int64_t nf = small_lds_floats(mg, sm, a.off, &a.scratch);
        a.nseg = 0;
        auto seg = [&](const float* src, int dst, int64_t n) {
            a.seg_src[a.nseg] = src;
            a.seg_dst[a.nseg] = dst;
            a.seg_len[a.nseg++] = (int)n;
        };
        // the top level's faces and b, then the coarser levels' faces
        for (int l = sm; l < L; ++l) {
            const Dims d = a.d[l];
            seg(mg.lev[l].bx, a.off[l][0], face_count(d, 0));
            seg(mg.lev[l].by, a.off[l][1], face_count(d, 1));
            seg(mg.lev[l].bz, a.off[l][2], face_count(d, 2));
            if (l == sm) {
                seg(B(sm), a.off[sm][4], ncell(mg.lev[sm]));
                a.nseg_first = a.nseg;
            }
        }
        a.done = done;
        a.cop = use_coarse_op(mg) ? static_cast<CGState*>(mg.scalars)->coarse_op : nullptr;
        a.cop_lds = -1;
        int64_t nlds = nf;
        if (a.cop) {
            // stage the coarsest level's matrix too when it fits the LDS
            // budget: its loads then overlap the faces' instead of costing
            // round trips of their own in the middle of the V-cycle
            const int64_t nc = ncell(mg.lev[L - 1]);
            const int64_t off = (nf + 3) & ~(int64_t)3;  // 16-byte aligned
            if ((off + 2 * nc * nc) * (int64_t)sizeof(float) <= kSmallMaxBytes) {
                a.cop_lds = (int)off;
                seg(reinterpret_cast<const float*>(a.cop), (int)off, 2 * nc * nc);
                nlds = off + 2 * nc * nc;
            }
        }
        const size_t lds = (size_t)nlds * sizeof(float);
        if (sm == 0) {
            a.p_top = z0;
            k_small_vcycle<ST><<<1, kSmallThreads, lds, s>>>(a);
        } else {
            a.p_top = mg.lev[sm].p;
            k_small_vcycle<float><<<1, kSmallThreads, lds, s>>>(a);
        }
        PF_CHECK_LAUNCH();
    } else {
        // coarsest level out of HBM: sweeps, then mean removal
        const int l = L - 1;
        int rc = smooth(l, mg.coarsest, false, true);
        if (rc) return rc;
        CGState* st = static_cast<CGState*>(mg.scalars);
        const int64_t n = ncell(mg.lev[l]);
        // p is ST; the mean is taken in FP64 and the shifted value stored back
        k_part_sum<ST><<<red_blocks(n), RT, 0, s>>>(cur[l], n, st->part);
        PF_CHECK_LAUNCH();
        k_comb<<<1, RT, 0, s>>>(st, C_MEAN, (int)red_blocks(n), n, &st->mu_x, 0.0);
        PF_CHECK_LAUNCH();
        k_shift<ST, ST><<<red_blocks(n), RT, 0, s>>>(cur[l], cur[l], n, &st->mu_x);
        PF_CHECK_LAUNCH();
    }
    for (int l = G - 1; l >= 0; --l) {
        const pf_mg_level& lv = mg.lev[l];
        const pf_mg_level& cv = mg.lev[l + 1];
        if (level_fused(mg, l, G)) {
            ST* out = other(l, cur[l]);
            const int rc =
                (l + 1 == sm)
                    ? launch_level<ST, false, float>(ldims(lv), ldims(cv), mg.post, Lv<ST>::fx(lv), Lv<ST>::fy(lv),
                                                     Lv<ST>::fz(lv), B(l), cur[l], cv.p, out, nullptr, mg.omega,
                                                     done, s)
                    : launch_level<ST, false, ST>(ldims(lv), ldims(cv), mg.post, Lv<ST>::fx(lv), Lv<ST>::fy(lv),
                                                  Lv<ST>::fz(lv), B(l), cur[l], cur[l + 1], out, nullptr,
                                                  mg.omega, done, s);
            if (rc) return rc;
            cur[l] = out;
            continue;
        }
        // the first post-sweep forms the prolongation itself when the last
        // post-sweep then lands where it must (level 0: in z0)
        if (fused_smoother(mg, l) && mg.post > 0 &&
            (l > 0 || ((mg.post % 2) ? other(0, cur[0]) : cur[0]) == z0)) {
            const Dims d = ldims(lv);
            ST* out = other(l, cur[l]);
            const float* bsc = l == 0 ? bscale : nullptr;
            int rc = (l + 1 == sm)
                         ? launch_full_sweep<ST, false, true, float>(d, Lv<ST>::fx(lv), Lv<ST>::fy(lv),
                                                                           Lv<ST>::fz(lv), cur[l], out, B(l), 1,
                                                                           mg.omega, done, bsc, s, cv.p, ldims(cv))
                               : launch_full_sweep<ST, false, true, ST>(d, Lv<ST>::fx(lv), Lv<ST>::fy(lv),
                                                                        Lv<ST>::fz(lv), cur[l], out, B(l), 1,
                                                                        mg.omega, done, bsc, s, cur[l + 1],
                                                                        ldims(cv));
            if (rc) return rc;
            cur[l] = out;
            rc = smooth(l, mg.post - 1, true, false);
            if (rc) return rc;
            continue;
        }
        // prolongation target: in place, except on level 0 where the last
        // post-sweep must land in z0
        ST* dst = cur[l];
        if (l == 0) dst = (fused_smoother(mg, 0) && (mg.post % 2)) ? second_buffer<ST>(mg, 0) : z0;
        const unsigned grid = cdiv(ncell(lv), RT);
        if (l + 1 == sm)
            k_prolong<ST, float><<<grid, RT, 0, s>>>(ldims(lv), ldims(cv), cur[l], dst, cv.p, done);
        else
            k_prolong<ST, ST><<<grid, RT, 0, s>>>(ldims(lv), ldims(cv), cur[l], dst, cur[l + 1], done);
        PF_CHECK_LAUNCH();
        cur[l] = dst;
        int rc = smooth(l, mg.post, true, false);
        if (rc) return rc;
    }
    return PF_OK;
}

// Host-side state of one solver (caller-owned through pf_mgpcg_state_*):
// the pinned int the residual tests publish to (the host polls it, no
// events), and the HIP graphs captured for this solver's buffers.  Bound to the device current at
// creation.
struct SolveGraphs;
struct SolverState {
    int device = -1;
    int* flag = nullptr;    // HW_WORDS pinned ints (host address)
    int* flag_d = nullptr;  // the same memory as the device addresses it
    int solves = 0;         // results published so far
    std::vector<SolveGraphs>* graphs = nullptr;
};

static void release_state(SolverState& st);

static int init_state(SolverState& st) {
    PF_CHECK_HIP(hipGetDevice(&st.device));
    // coherent: the device's stores reach host memory without cache flushes
    PF_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&st.flag), HW_WORDS * sizeof(int),
                               hipHostMallocMapped | hipHostMallocCoherent));
    for (int w = 0; w < HW_WORDS; ++w) st.flag[w] = 0;
    PF_CHECK_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&st.flag_d), st.flag, 0));
    return PF_OK;
}

// The V-cycle's right-hand-side scale of a solve that reads z scaled.
template <typename ST>
static const float* vcycle_scale(const CGK<ST>& a) {
    return a.zscale ? &a.st->vscale : nullptr;
}

// Kernels of one FPCG iteration; `flag` (the device address of pinned host
// memory) receives the done flag from the residual test's combine kernel.
// Every kernel returns at once when the flag is set, so an iteration
// enqueued past convergence does nothing.
// The residual test rides in the V-cycle's first kernel when that is the
// fused zero-start sweep of an HBM level 0.
static bool fold_residual_test(const pf_mg& mg) {
    return mg.small_from > 0 && mg.pre > 0 && fused_smoother(mg, 0);
}

// The inner products after the V-cycle ride on its last level-0 sweep when
// that is a plain fused sweep (at least two post-sweeps on an HBM level 0)
// whose workgroups fit the partials array.
static bool fold_dots(const pf_mg& mg) {
    return !(mg.options & PF_MG_SEPARATE_DOTS) && mg.small_from > 0 && mg.post >= 2 && fused_smoother(mg, 0) &&
           full_sweep_blocks(ldims(mg.lev[0])) <= MAXB;
}

template <typename ST>
static int enqueue_iteration(const pf_mg& mg, const CGK<ST>& a, ST* z, const int* done, int* flag,
                             unsigned nb, hipStream_t s) {
    CGState* st = a.st;
    k_cg_apply<ST><<<a.np_apply, WT, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    k_cg_update<ST><<<nb, RT, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    const UpdateFold uf{st, (int)nb, flag};
    const bool fold = fold_residual_test(mg);
    if (!fold) {
        k_comb<<<1, RT, 0, s>>>(st, C_UPDATE, (int)nb, a.n, nullptr, 0.0, flag);
        PF_CHECK_LAUNCH();
    }
    const DotFold df{a.Ap, a.zscale, st};
    const bool dots = fold_dots(mg);
    bool dot_done = false;
    int rc = vcycle<ST>(mg, a.r, z, done, vcycle_scale(a), s, fold ? &uf : nullptr, dots ? &df : nullptr,
                        &dot_done);
    if (rc) return rc;
    PF_REQUIRE(dot_done == dots, "pf_mgpcg_solve: the V-cycle's last level-0 kernel is not a plain fused sweep");
    if (!dots) {
        k_cg_beta<ST><<<a.np_beta, WT, 0, s>>>(a);
        PF_CHECK_LAUNCH();
    }
    return PF_OK;
}

// First preconditioner application (poisson.py:378-380); the first search
// direction is formed inside the first iteration's k_cg_apply (beta = 0).
template <typename ST>
static int enqueue_prologue(const pf_mg& mg, const CGK<ST>& a, ST* z, const int* done, unsigned nb,
                            hipStream_t s) {
    int rc = vcycle<ST>(mg, a.r, z, done, vcycle_scale(a), s);
    if (rc) return rc;
    k_cg_first<ST><<<nb, RT, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    k_comb<<<1, RT, 0, s>>>(a.st, C_FIRST, (int)nb, a.n, nullptr, 0.0);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

// _remove_mean(x) (poisson.py:408), then the result for the host
static int enqueue_epilogue(CGState* st, float* x, int64_t n, unsigned nb, int* words, hipStream_t s) {
    k_part_sum<float><<<nb, RT, 0, s>>>(x, n, st->part);
    PF_CHECK_LAUNCH();
    k_comb<<<1, RT, 0, s>>>(st, C_MEAN, (int)nb, n, &st->mu_x, 0.0);
    PF_CHECK_LAUNCH();
    k_shift<float, float><<<nb, RT, 0, s>>>(x, x, n, &st->mu_x);
    PF_CHECK_LAUNCH();
    k_publish_result<<<1, 64, 0, s>>>(st, words);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

// HIP graphs of the prologue, of one iteration (two instances that differ
// only in which search-direction buffer they write) and of an odd and an
// even iteration together, cached per solver layout.  A solve issues one
// graph launch per two iterations instead of ~20 kernel launches each.
// pair: an odd iteration and the even one after it in one graph; the GPU
// spends ~8 us between two graphs, none between two kernels of one graph
struct SolveGraphs {
    std::vector<uint64_t> key;
    hipGraphExec_t pro = nullptr;
    hipGraphExec_t it[2] = {nullptr, nullptr};
    hipGraphExec_t pair = nullptr;
    hipGraphExec_t epi = nullptr;  // mean removal and the result for the host
};

static void destroy_graphs(SolveGraphs& g) {
    if (g.pro) (void)hipGraphExecDestroy(g.pro);
    for (auto& e : g.it)
        if (e) (void)hipGraphExecDestroy(e);
    if (g.pair) (void)hipGraphExecDestroy(g.pair);
    if (g.epi) (void)hipGraphExecDestroy(g.epi);
    g.pro = g.it[0] = g.it[1] = g.pair = g.epi = nullptr;
}

template <typename F>
static int capture(hipStream_t cs, F&& enqueue, hipGraphExec_t* out) {
    PF_CHECK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
    const int rc = enqueue(cs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return set_error((int)e, "hipStreamEndCapture: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return set_error((int)ei, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    return PF_OK;
}

// Iteration `slot` (= iteration number & 1) writes its search direction to
// mg.q (slot 0) or mg.rhs (slot 1) and reads the other one.
template <typename ST>
static CGK<ST> slot_args(const CGK<ST>& a, const pf_mg& mg, int slot) {
    CGK<ST> b = a;
    b.q = slot ? mg.rhs : mg.q;
    b.q_old = slot ? mg.q : mg.rhs;
    b.slot = slot;
    return b;
}

// Everything a captured graph bakes in: the hierarchy's sizes, scalars and
// buffer addresses, the solve's vectors, the flag slot and the launch shape.
template <typename ST>
static std::vector<uint64_t> graph_key(const pf_mg& mg, const CGK<ST>& a, const ST* z, const int* flag,
                                       unsigned nb) {
    std::vector<uint64_t> k;
    auto ptr = [&](const void* p) { k.push_back((uint64_t)(uintptr_t)p); };
    uint32_t om;
    std::memcpy(&om, &mg.omega, sizeof(om));
    k.insert(k.end(), {(uint64_t)mg.nlev, (uint64_t)mg.pre, (uint64_t)mg.post, (uint64_t)mg.coarsest, om,
                       (uint64_t)mg.precision, (uint64_t)mg.small_from, (uint64_t)mg.options,
                       (uint64_t)sizeof(ST), (uint64_t)nb});
    for (int l = 0; l < mg.nlev; ++l) {
        const pf_mg_level& v = mg.lev[l];
        k.insert(k.end(), {(uint64_t)v.nx, (uint64_t)v.ny, (uint64_t)v.nz});
        for (const void* p : {(const void*)v.bx, (const void*)v.by, (const void*)v.bz, (const void*)v.diag,
                              (const void*)v.p, (const void*)v.b, (const void*)v.r, (const void*)v.bx_h,
                              (const void*)v.by_h, (const void*)v.bz_h, (const void*)v.p_h})
            ptr(p);
    }
    for (const void* p : {(const void*)mg.rhs, (const void*)mg.res, (const void*)mg.z, (const void*)mg.q,
                          (const void*)mg.Ap, (const void*)mg.scalars, (const void*)a.b, (const void*)a.x,
                          (const void*)z, (const void*)flag})
        ptr(p);
    return k;
}

template <typename ST>
static int get_graphs(SolverState& st, const pf_mg& mg, const CGK<ST>& a, ST* z, const int* done,
                      unsigned nb, SolveGraphs*& out) {
    std::vector<SolveGraphs>& cache = *st.graphs;
    std::vector<uint64_t> key = graph_key<ST>(mg, a, z, st.flag, nb);
    for (auto& c : cache)
        if (c.key == key) {
            out = &c;
            return PF_OK;
        }
    int rc = ensure_small_lds();
    if (rc) return rc;
    SolveGraphs g;
    g.key = std::move(key);
    hipStream_t cs = nullptr;
    PF_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    rc = capture(cs, [&](hipStream_t t) { return enqueue_prologue<ST>(mg, a, z, done, nb, t); }, &g.pro);
    for (int q = 0; q < 2 && !rc; ++q) {
        const CGK<ST> aq = slot_args(a, mg, q);
        rc = capture(cs, [&](hipStream_t t) { return enqueue_iteration<ST>(mg, aq, z, done, st.flag_d, nb, t); },
                     &g.it[q]);
    }
    if (!rc) {
        const CGK<ST> a1 = slot_args(a, mg, 1), a0 = slot_args(a, mg, 0);
        rc = capture(
            cs,
            [&](hipStream_t t) {
                const int r1 = enqueue_iteration<ST>(mg, a1, z, done, st.flag_d, nb, t);
                return r1 ? r1 : enqueue_iteration<ST>(mg, a0, z, done, st.flag_d, nb, t);
            },
            &g.pair);
    }
    if (!rc)
        rc = capture(
            cs, [&](hipStream_t t) { return enqueue_epilogue(a.st, a.x, a.n, nb, st.flag_d, t); }, &g.epi);
    (void)hipStreamDestroy(cs);
    if (rc) {
        destroy_graphs(g);
        return rc;
    }
    // a solver is normally solved with one set of vectors; keep a few
    if (cache.size() >= 4) {
        destroy_graphs(cache.front());
        cache.erase(cache.begin());
    }
    cache.push_back(std::move(g));
    out = &cache.back();
    return PF_OK;
}

static void release_state(SolverState& st) {
    if (st.graphs) {
        for (auto& g : *st.graphs) destroy_graphs(g);
        delete st.graphs;
        st.graphs = nullptr;
    }
    if (st.flag) (void)hipHostFree(st.flag);
    st.flag = st.flag_d = nullptr;
}

static bool use_graphs() {
    static const bool on = [] {
        const char* e = std::getenv("PFFLIP_SOLVER_GRAPHS");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Spin until the published counter (see publish_done) reaches `seq`; *stop
// = the solve is done.  The stream running dry without that publication is
// an error, so the loop cannot spin forever.
static int wait_published(const int* flag, int seq, hipStream_t s, bool* stop) {
    const volatile int* f = flag;
    for (unsigned spin = 0;; ++spin) {
        const int v = *f;
        if ((v >> 1) >= seq) {
            *stop = (v & 1) != 0;
            return PF_OK;
        }
        if ((spin & 255) == 255) {
            const hipError_t q = hipStreamQuery(s);
            if (q == hipSuccess) {
                const int w = *f;
                if ((w >> 1) >= seq) {
                    *stop = (w & 1) != 0;
                    return PF_OK;
                }
                return set_error(PF_ERR_INVALID, "pf_mgpcg_solve: residual test %d was never published", seq);
            }
            if (q != hipErrorNotReady)
                return set_error((int)q, "pf_mgpcg_solve: %s", hipGetErrorString(q));
        }
        __builtin_ia32_pause();
    }
}

// Spin until k_publish_result stored `number` (the stream running dry
// without it is an error, so the loop cannot spin forever); the result
// words are read after it (x86 keeps loads in order).
static int wait_result(const int* words, int number, hipStream_t s) {
    const volatile int* f = words + HW_RESULT_NO;
    for (unsigned spin = 0;; ++spin) {
        if (*f == number) break;
        if ((spin & 255) == 255) {
            const hipError_t q = hipStreamQuery(s);
            if (q == hipSuccess) {
                if (*f == number) break;
                return set_error(PF_ERR_INVALID, "pf_mgpcg_solve: the result was never published");
            }
            if (q != hipErrorNotReady) return set_error((int)q, "pf_mgpcg_solve: %s", hipGetErrorString(q));
        }
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return PF_OK;
}

template <typename ST>
static int solve_impl(const pf_mg& mg, const float* b, float* x, double tol, int max_iter,
                      int* it_out, double* rel_out, hipStream_t s) {
    CGState* st = static_cast<CGState*>(mg.scalars);
    const pf_mg_level& l0 = mg.lev[0];
    const Dims d = ldims(l0);
    const int64_t n = ncell(l0);
    ST* z = (sizeof(ST) == 4) ? reinterpret_cast<ST*>(mg.z) : Lv<ST>::p(l0);
    CGK<ST> a;
    a.d = d;
    a.n = n;
    a.bx = l0.bx;
    a.by = l0.by;
    a.bz = l0.bz;
    a.b = b;
    a.rhs = mg.rhs;
    a.x = x;
    a.r = mg.res;
    a.q = mg.q;  // zeroed by k_cg_init: iteration 1 (slot 1) reads it as q_old
    a.q_old = mg.rhs;
    a.Ap = mg.Ap;
    a.z = z;
    // 16-bit iterates: scale the V-cycle's input (not for a whole hierarchy
    // in LDS, which computes in FP32 and rounds only its output)
    a.zscale = (sizeof(ST) == 2 && mg.small_from > 0) ? &st->vscale_inv : nullptr;
    a.st = st;
    a.tol = tol;
    const unsigned nb = red_blocks(n);
    a.slot = 1;
    a.np_apply = a.np_beta = (int)std::min<int64_t>(WB, (n + WT - 1) / WT);
    if (fold_dots(mg)) a.np_beta = (int)full_sweep_blocks(d);  // one partial per level-0 tile
    const int* done = &st->done;

    k_part_sum<float><<<nb, RT, 0, s>>>(b, n, st->part);
    PF_CHECK_LAUNCH();
    k_comb<<<1, RT, 0, s>>>(st, C_MEAN_RESET, (int)nb, n, &st->mu, 0.0);
    PF_CHECK_LAUNCH();
    k_cg_init<ST><<<nb, RT, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    // b == 0 after mean removal (poisson.py:366-368): C_INIT sets done = 2
    // and rel = 0, x stays k_cg_init's zeros, every later kernel exits on the
    // flag and the first residual test stops the host loop, so the result
    // is (0, 0) without a synchronisation here
    k_comb<<<1, RT, 0, s>>>(st, C_INIT, (int)nb, n, nullptr, tol);
    PF_CHECK_LAUNCH();
    // the caller's solver state (graphs, flags), or a temporary one for this
    // call (no graphs)
    SolverState local;
    struct LocalGuard {
        SolverState& s;
        ~LocalGuard() { release_state(s); }
    } guard{local};
    SolverState* hp = static_cast<SolverState*>(mg.host_state);
    int rc = PF_OK;
    if (hp) {
        int dev = 0;
        PF_CHECK_HIP(hipGetDevice(&dev));
        PF_REQUIRE(dev == hp->device, "pf_mgpcg_solve: solver state was created on device %d, current is %d",
                   hp->device, dev);
    } else {
        rc = init_state(local);
        if (rc) return rc;
        hp = &local;
    }
    // the previous solve's last publication must not be read as this one's
    // (its kernels have all completed: it ended with a stream synchronize)
    static_cast<volatile int*>(hp->flag)[0] = 0;
    SolveGraphs* gr = nullptr;
    if (use_graphs() && max_iter > 0 && hp->graphs) {
        rc = get_graphs<ST>(*hp, mg, a, z, done, nb, gr);
        if (rc) return rc;
    }
    if (gr) {
        PF_CHECK_HIP(hipGraphLaunch(gr->pro, s));
    } else {
        rc = enqueue_prologue<ST>(mg, a, z, done, nb, s);
        if (rc) return rc;
    }
    for (int it = 1; it <= max_iter;) {
        const int slot = it & 1;
        // graphs: iterations in (odd, even) pairs, one graph each (single
        // iterations without graphs and for an odd max_iter's last).  Then
        // wait for the last launched iteration's residual test: it opens
        // that iteration's first kernel, so the rest of the iteration (about
        // 200 us at 128^3) is still queued while the host launches the next
        // one, and after convergence at most the iteration that found it and
        // one more run (as early exits).
        int step = 1;
        if (gr) {
            if (slot == 1 && it < max_iter) {
                PF_CHECK_HIP(hipGraphLaunch(gr->pair, s));
                step = 2;
            } else {
                PF_CHECK_HIP(hipGraphLaunch(gr->it[slot], s));
            }
        } else {
            rc = enqueue_iteration<ST>(mg, slot_args(a, mg, slot), z, done, hp->flag_d, nb, s);
            if (rc) return rc;
        }
        bool stop = false;
        rc = wait_published(hp->flag, it + step - 1, s, &stop);
        if (rc) return rc;
        if (stop) break;
        it += step;
    }
    // _remove_mean(x) (poisson.py:408) and the result, published under the
    // solve's number
    const int number = ++hp->solves;
    static_cast<volatile int*>(hp->flag)[HW_WANT] = number;
    if (gr) {
        PF_CHECK_HIP(hipGraphLaunch(gr->epi, s));
    } else {
        rc = enqueue_epilogue(st, x, n, nb, hp->flag_d, s);
        if (rc) return rc;
    }
    rc = wait_result(hp->flag, number, s);
    if (rc) return rc;
    const volatile int* w = hp->flag;
    *it_out = w[HW_IT];
    *rel_out = __builtin_bit_cast(double, *reinterpret_cast<const volatile unsigned long long*>(w + HW_REL));
    return PF_OK;
}

template <typename ST>
static int precondition_impl(const pf_mg& mg, const float* r, float* z, hipStream_t s) {
    CGState* st = static_cast<CGState*>(mg.scalars);
    const int64_t n = ncell(mg.lev[0]);
    ST* zt = (sizeof(ST) == 4) ? reinterpret_cast<ST*>(z) : Lv<ST>::p(mg.lev[0]);
    int rc = vcycle<ST>(mg, r, zt, nullptr, nullptr, s);
    if (rc) return rc;
    k_part_sum<ST><<<red_blocks(n), RT, 0, s>>>(zt, n, st->part);
    PF_CHECK_LAUNCH();
    k_comb<<<1, RT, 0, s>>>(st, C_MEAN, (int)red_blocks(n), n, &st->mu_x, 0.0);
    PF_CHECK_LAUNCH();
    k_shift<ST, float><<<red_blocks(n), RT, 0, s>>>(zt, z, n, &st->mu_x);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

static int check_mg(const pf_mg* mg) {
    PF_REQUIRE(mg, "null pf_mg");
    PF_REQUIRE(mg->nlev >= 1 && mg->nlev <= PF_MAX_LEVELS, "pf_mg: bad level count %d", mg->nlev);
    PF_REQUIRE(mg->precision >= 0 && mg->precision <= 2, "pf_mg: bad precision %d", mg->precision);
    PF_REQUIRE(mg->pre >= 0 && mg->post >= 0 && mg->coarsest >= 0, "pf_mg: negative sweep count");
    PF_REQUIRE(mg->scalars, "pf_mg: null scalar workspace");
    PF_REQUIRE(mg->small_from >= 0 && mg->small_from <= mg->nlev, "pf_mg: call pf_mg_setup first");
    for (int l = 0; l < mg->nlev; ++l) {
        const pf_mg_level& v = mg->lev[l];
        PF_REQUIRE(v.nx > 0 && v.ny > 0 && v.nz > 0, "pf_mg: level %d has bad dims", l);
        // the kernels index cells and faces with 32-bit ints
        PF_REQUIRE((int64_t)(v.nx + 1) * (v.ny + 1) * (v.nz + 1) < ((int64_t)1 << 31), "pf_mg: level %d too large", l);
        PF_REQUIRE(v.bx && v.by && v.bz && v.b, "pf_mg: level %d missing arrays", l);
        if (l > 0) {
            const pf_mg_level& f = mg->lev[l - 1];
            PF_REQUIRE(f.nx == 2 * v.nx && f.ny == 2 * v.ny && f.nz == 2 * v.nz,
                       "pf_mg: level %d is not the 2x coarsening of level %d", l, l - 1);
        }
        if (mg->precision != 0 && l < mg->small_from)
            PF_REQUIRE(v.bx_h && v.by_h && v.bz_h && v.p_h, "pf_mg: level %d missing 16-bit arrays", l);
        // level 0's 16-bit iterate holds the preconditioned vector z
        if (mg->precision != 0 && l == 0) PF_REQUIRE(v.p_h, "pf_mg: level 0 missing its 16-bit iterate");
        if (l > 0 && (mg->precision == 0 || l >= mg->small_from)) PF_REQUIRE(v.p, "pf_mg: level %d missing p", l);
        // the fused smoother's second iterate buffer on the HBM levels
        if (l < mg->small_from) {
            PF_REQUIRE(v.r, "pf_mg: level %d missing r (second smoother buffer)", l);
            if (l == 0 && mg->precision == 0) PF_REQUIRE(v.p, "pf_mg: level 0 missing p (second smoother buffer)");
        }
    }
    return PF_OK;
}

}  // namespace pf

using namespace pf;

PF_API size_t pf_mg_scalar_bytes(void) { return sizeof(CGState); }

PF_API int pf_mg_setup(pf_mg* mg) {
    PF_REQUIRE(mg, "pf_mg_setup: null pf_mg");
    PF_REQUIRE(mg->nlev >= 1 && mg->nlev <= PF_MAX_LEVELS, "pf_mg_setup: bad level count");
    int sm = mg->nlev;
    for (int l = mg->nlev - 1; l >= 0; --l) {
        const int64_t nc = (int64_t)mg->lev[l].nx * mg->lev[l].ny * mg->lev[l].nz;
        const int64_t bytes = small_lds_floats(*mg, l, nullptr, nullptr) * (int64_t)sizeof(float);
        const Dims d{mg->lev[l].nx, mg->lev[l].ny, mg->lev[l].nz};
        if (nc <= kSmallMaxCells && bytes <= kSmallMaxBytes && colour_space(d) <= kSmallThreads * kSlots)
            sm = l;
        else
            break;
    }
    mg->small_from = sm;
    return PF_OK;
}

PF_API int pf_mg_build(const pf_mg* mg, const float* bx, const float* by, const float* bz,
                       pf_stream_t stream) {
    int rc = check_mg(mg);
    if (rc) return rc;
    PF_REQUIRE(bx && by && bz, "pf_mg_build: null coefficients");
    hipStream_t s = S(stream);
    const pf_mg_level& l0 = mg->lev[0];
    const Dims d0 = ldims(l0);
    const float* src[3] = {bx, by, bz};
    float* dst[3] = {l0.bx, l0.by, l0.bz};
    for (int q = 0; q < 3; ++q)
        if (src[q] != dst[q])
            PF_CHECK_HIP(hipMemcpyAsync(dst[q], src[q], face_count(d0, q) * sizeof(float),
                                        hipMemcpyDeviceToDevice, s));
    for (int l = 0; l < mg->nlev; ++l) {
        const pf_mg_level& v = mg->lev[l];
        const Dims d = ldims(v);
        if (l > 0) {
            const pf_mg_level& f = mg->lev[l - 1];
            const int64_t n = (int64_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
            k_galerkin<<<cdiv(n, RT), RT, 0, s>>>(ldims(f), d, f.bx, f.by, f.bz, v.bx, v.by, v.bz);
            PF_CHECK_LAUNCH();
        }
        const int64_t n = (int64_t)(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
        const bool copy = mg->precision != 0 && l < mg->small_from;
        if (mg->precision == 2)
            k_diag_copy<fp16s><<<cdiv(n, RT), RT, 0, s>>>(
                d, v.bx, v.by, v.bz, v.diag, copy ? reinterpret_cast<fp16s*>(v.bx_h) : nullptr,
                reinterpret_cast<fp16s*>(v.by_h), reinterpret_cast<fp16s*>(v.bz_h));
        else
            k_diag_copy<bf16s><<<cdiv(n, RT), RT, 0, s>>>(
                d, v.bx, v.by, v.bz, v.diag, copy ? reinterpret_cast<bf16s*>(v.bx_h) : nullptr,
                reinterpret_cast<bf16s*>(v.by_h), reinterpret_cast<bf16s*>(v.bz_h));
        PF_CHECK_LAUNCH();
    }
    if (use_coarse_op(*mg)) {
        const pf_mg_level& c = mg->lev[mg->nlev - 1];
        const int n = c.nx * c.ny * c.nz;
        const int threads = ((n + 63) / 64) * 64;
        k_coarse_op<<<n, threads, 0, s>>>(ldims(c), c.bx, c.by, c.bz, mg->coarsest, mg->omega,
                                          static_cast<CGState*>(mg->scalars)->coarse_op);
        PF_CHECK_LAUNCH();
    }
    return PF_OK;
}

PF_API int pf_mg_precondition(const pf_mg* mg, const float* r, float* z, pf_stream_t stream) {
    int rc = check_mg(mg);
    if (rc) return rc;
    PF_REQUIRE(r && z, "pf_mg_precondition: null vector");
    switch (mg->precision) {
        case 0: return precondition_impl<float>(*mg, r, z, S(stream));
        case 1: return precondition_impl<bf16s>(*mg, r, z, S(stream));
        default: return precondition_impl<fp16s>(*mg, r, z, S(stream));
    }
}

// One V-cycle from a zero guess, without the final mean removal of
// pf_mg_precondition.  FP32 hierarchies only: the 16-bit V-cycles keep their
// level-0 iterate in 16-bit storage with a scale that pf_mg_precondition
// and pf_mgpcg_solve undo.
PF_API int pf_mg_vcycle(const pf_mg* mg, const float* r, float* z, pf_stream_t stream) {
    int rc = check_mg(mg);
    if (rc) return rc;
    PF_REQUIRE(r && z, "pf_mg_vcycle: null vector");
    PF_REQUIRE(mg->precision == 0, "pf_mg_vcycle: FP32 hierarchies only (precision %d)", mg->precision);
    // level 0 ping-pongs between z and the hierarchy's own p
    PF_REQUIRE((const void*)z != (const void*)r && (const void*)z != (const void*)mg->lev[0].p &&
                   (const void*)r != (const void*)mg->lev[0].p,
               "pf_mg_vcycle: r, z and the level-0 p buffer must be distinct");
    return vcycle<float>(*mg, r, z, nullptr, nullptr, S(stream));
}

PF_API int pf_mgpcg_state_create(void** out) {
    PF_REQUIRE(out, "pf_mgpcg_state_create: null argument");
    *out = nullptr;
    SolverState* st = new (std::nothrow) SolverState();
    PF_REQUIRE(st, "pf_mgpcg_state_create: out of host memory");
    int rc = init_state(*st);
    if (!rc) {
        st->graphs = new (std::nothrow) std::vector<SolveGraphs>();
        if (!st->graphs) rc = set_error(PF_ERR_INVALID, "pf_mgpcg_state_create: out of host memory");
    }
    if (rc) {
        release_state(*st);
        delete st;
        return rc;
    }
    *out = st;
    return PF_OK;
}

PF_API int pf_mgpcg_state_destroy(void* state) {
    if (!state) return PF_OK;
    SolverState* st = static_cast<SolverState*>(state);
    int cur = 0;
    const bool switched = hipGetDevice(&cur) == hipSuccess && cur != st->device;
    if (switched) (void)hipSetDevice(st->device);
    release_state(*st);
    if (switched) (void)hipSetDevice(cur);
    delete st;
    return PF_OK;
}

PF_API int pf_mgpcg_solve(const pf_mg* mg, const float* b, float* x, double tol, int32_t max_iter,
                          int32_t* it_out, double* rel_out, pf_stream_t stream) {
    int rc = check_mg(mg);
    if (rc) return rc;
    PF_REQUIRE(b && x && it_out && rel_out, "pf_mgpcg_solve: null argument");
    PF_REQUIRE(mg->rhs && mg->res && mg->q && mg->Ap, "pf_mgpcg_solve: missing CG vectors");
    PF_REQUIRE(mg->precision != 0 || mg->z, "pf_mgpcg_solve: missing z");
    PF_REQUIRE(max_iter >= 0, "pf_mgpcg_solve: negative max_iter");
    switch (mg->precision) {
        case 0: return solve_impl<float>(*mg, b, x, tol, max_iter, it_out, rel_out, S(stream));
        case 1: return solve_impl<bf16s>(*mg, b, x, tol, max_iter, it_out, rel_out, S(stream));
        default: return solve_impl<fp16s>(*mg, b, x, tol, max_iter, it_out, rel_out, S(stream));
    }
}

template <typename ST>
static int time_smoother_impl(const pf_mg& mg, int level, int reps, float* ms_out, hipStream_t s) {
    const pf_mg_level& lv = mg.lev[level];
    const Dims d = ldims(lv);
    ST* pa = level == 0 ? (sizeof(ST) == 4 ? reinterpret_cast<ST*>(mg.z) : Lv<ST>::p(lv)) : Lv<ST>::p(lv);
    ST* pb = second_buffer<ST>(mg, level);
    const float* b = level == 0 ? mg.res : lv.b;
    const bool fused = fused_smoother(mg, level);
    const int64_t nthr = (int64_t)d.nx * d.ny * ((d.nz + 1) / 2);
    const unsigned grid = cdiv(nthr, RT);
    hipEvent_t e0, e1;
    PF_CHECK_HIP(hipEventCreate(&e0));
    PF_CHECK_HIP(hipEventCreate(&e1));
    auto run = [&]() -> int {
        for (int r = 0; r < reps; ++r) {
            if (fused) {
                const int rc = launch_full_sweep<ST, false>(d, Lv<ST>::fx(lv), Lv<ST>::fy(lv), Lv<ST>::fz(lv),
                                                            r % 2 ? pb : pa, r % 2 ? pa : pb, b, 0, mg.omega,
                                                            nullptr, nullptr, s);
                if (rc) return rc;
            } else {
                for (int c = 0; c < 2; ++c) {
                    k_rbgs<ST, false><<<grid, RT, 0, s>>>(d, Lv<ST>::fx(lv), Lv<ST>::fy(lv), Lv<ST>::fz(lv), pa,
                                                          b, c, mg.omega, nullptr, nullptr);
                    PF_CHECK_LAUNCH();
                }
            }
        }
        return PF_OK;
    };
    int rc = run();  // warm-up
    if (!rc) {
        PF_CHECK_HIP(hipEventRecord(e0, s));
        rc = run();
    }
    if (!rc) {
        PF_CHECK_HIP(hipEventRecord(e1, s));
        PF_CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PF_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_out = ms / (fused ? (float)reps : 2.0f * reps);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

PF_API int pf_mg_time_smoother(const pf_mg* mg, int32_t level, int32_t reps, float* ms_out,
                               pf_stream_t stream) {
    int rc = check_mg(mg);
    if (rc) return rc;
    PF_REQUIRE(ms_out && reps > 0, "pf_mg_time_smoother: bad arguments");
    PF_REQUIRE(level >= 0 && level < mg->nlev && level < mg->small_from,
               "pf_mg_time_smoother: level %d is not smoothed out of HBM", level);
    PF_REQUIRE(level > 0 || mg->precision != 0 || mg->z, "pf_mg_time_smoother: missing z");
    switch (mg->precision) {
        case 0: return time_smoother_impl<float>(*mg, level, reps, ms_out, S(stream));
        case 1: return time_smoother_impl<bf16s>(*mg, level, reps, ms_out, S(stream));
        default: return time_smoother_impl<fp16s>(*mg, level, reps, ms_out, S(stream));
    }
}

PF_API int pf_apply_A(int nx, int ny, int nz, const float* p, const float* bx, const float* by,
                      const float* bz, const float* diag, float* out, pf_stream_t stream) {
    PF_REQUIRE(nx > 0 && ny > 0 && nz > 0, "pf_apply_A: bad grid");
    PF_REQUIRE(p && bx && by && bz && out, "pf_apply_A: null argument");
    const int64_t n = (int64_t)nx * ny * nz;
    k_apply_A<<<cdiv(n, RT), RT, 0, S(stream)>>>(Dims{nx, ny, nz}, p, bx, by, bz, diag, out);
    PF_CHECK_LAUNCH();
    return PF_OK;
}
