// Multiresolution sparse block grid (MSBG): addressing, the 2:1 variable-
// coefficient Poisson operator and the aggregation multigrid transfers
// (pfflip/msbg.py, pfflip/mr_operator.py, pfflip/mr_multigrid.py).
//
// Blocks have one geometric size; a level-l block holds (n0 >> l)^3 cells.
// Every lookup goes through finest-level integer coordinates: block =
// F >> log2(n0), cell = (F & (n0 - 1)) >> l.  Each cell stores its three
// low-side faces; the face between two cells is read from the cell on its
// high side, so both rows see one coefficient and the operator is symmetric
// across resolution jumps.
//
// The red-black smoother reads one buffer and writes another: every cell of
// the active colour is updated from the values before the half-sweep.  With
// the finest-coordinate colouring two same-colour cells can touch across a
// resolution jump; reading the old values there makes the half-sweep
// deterministic.  Restriction gathers each coarse cell's children in a fixed
// order instead of adding them with atomics.
//
// The cell-parallel operator kernels take their cell range through
// xcd_block: the 16 workgroups of a 16^3 block then run on one XCD and the
// +-z neighbour rows they read from each other stay in that XCD's L2.
#include "pf_msbg.hpp"

namespace pf {

int fill_u32(void* dst, int64_t n, uint32_t v, hipStream_t s);

// Calls f(neighbour address, T) for every coupling of cell `addr`, with
// T = beta * min(hA, hB)^2 / ((hA + hB) / 2), walking the faces in the
// reference's order (mr_operator.py:74-166); across a face to finer cells
// the 2x2 patch behind it.
template <class F>
__device__ inline void for_each_coupling_from(const MGrid& g, const float* __restrict__ bface, int addr,
                                              const int4 o, const BlockRef& br, F&& f) {
    const int la = o.w;
    const int ext = 1 << la;
    const float hA = g.h0 * (float)ext;
    for (int axis = 0; axis < 3; ++axis) {
        const int ex = comp(axis, 0), ey = comp(axis, 1), ez = comp(axis, 2);
        for (int side = 0; side < 2; ++side) {
            int fx, fy, fz;
            if (side == 0) {
                fx = o.x - ex;
                fy = o.y - ey;
                fz = o.z - ez;
            } else {
                fx = o.x + ex * ext;
                fy = o.y + ey * ext;
                fz = o.z + ez * ext;
            }
            const MCell nb = locate_near(g, br, fx, fy, fz);
            if (nb.addr < 0) continue;  // domain wall: no flux
            const float hB = g.h0 * (float)(1 << nb.level);
            if (nb.level >= la) {
                // same size or coarser: one neighbour
                float area = fminf(hA, hB);
                area = area * area;
                const float dist = 0.5f * (hA + hB);
                const float beta = side == 0 ? bface[addr * 3 + axis] : bface[nb.addr * 3 + axis];
                f(nb.addr, beta * area / dist);
            } else {
                // finer: the 2x2 patch across this face
                const int sub = 1 << nb.level;
                const float area = hB * hB;
                const float dist = 0.5f * (hA + hB);
                const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
                for (int u = 0; u < 2; ++u)
                    for (int v = 0; v < 2; ++v) {
                        const int du = u * sub, dv = v * sub;
                        const MCell nc = locate_near(g, br, fx + comp(a1, 0) * du + comp(a2, 0) * dv,
                                                     fy + comp(a1, 1) * du + comp(a2, 1) * dv,
                                                     fz + comp(a1, 2) * du + comp(a2, 2) * dv);
                        if (nc.addr < 0) continue;
                        const float beta = side == 0 ? bface[addr * 3 + axis] : bface[nc.addr * 3 + axis];
                        f(nc.addr, beta * area / dist);
                    }
            }
        }
    }
}

template <class F>
__device__ inline void for_each_coupling(const MGrid& g, const float* __restrict__ bface, int addr, F&& f) {
    BlockRef br;
    const int4 o = origin_ref(g, addr, br);
    for_each_coupling_from(g, bface, addr, o, br, f);
}

// sum_f T_f (p_A - p_B) over the couplings of cell `addr` (origin o, block
// br), or (want_diag) the sum of T_f
__device__ inline float apply_row_from(const MGrid& g, const float* __restrict__ p,
                                       const float* __restrict__ bface, int addr, const int4 o,
                                       const BlockRef& br, bool want_diag) {
    const float pa = want_diag ? 0.0f : p[addr];
    float acc = 0.0f, diag = 0.0f;
    for_each_coupling_from(g, bface, addr, o, br, [&](int nb, float t) {
        diag += t;
        if (!want_diag) acc += t * (pa - p[nb]);
    });
    return want_diag ? diag : acc;
}

__device__ inline float apply_row(const MGrid& g, const float* __restrict__ p, const float* __restrict__ bface,
                                  int addr, bool want_diag) {
    BlockRef br;
    const int4 o = origin_ref(g, addr, br);
    return apply_row_from(g, p, bface, addr, o, br, want_diag);
}

__global__ __launch_bounds__(256) void k_msbg_apply(MGrid g, const float* __restrict__ p,
                                                    const float* __restrict__ bface, float* __restrict__ out,
                                                    int want_diag) {
    const int a = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    out[a] = apply_row(g, p, bface, a, want_diag != 0);
}

__global__ __launch_bounds__(256) void k_msbg_residual(MGrid g, const float* __restrict__ p,
                                                       const float* __restrict__ b,
                                                       const float* __restrict__ bface, float* __restrict__ r) {
    const int a = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    r[a] = b[a] - apply_row(g, p, bface, a, false);
}

// Grids from this many cells up take the paired half-sweep: on the 1.4M-cell
// level of a 128^3 dam break it runs 19.6 us against 25.9; on levels of a
// few thousand cells, mostly small blocks, it was up to twice as slow.
constexpr int kMsbgPairMinCells = 1 << 17;

// one cell of the half-sweep: the weighted update if the cell has the
// active colour and diag > 0, else a copy
__device__ inline void rbgs_cell(const MGrid& g, const float* __restrict__ pin, float* __restrict__ pout,
                                 const float* __restrict__ b, const float* __restrict__ bface,
                                 const float* __restrict__ diag, int colour, float omega, int a, const int4 o,
                                 const BlockRef& br) {
    float v = pin[a];
    if ((((o.x >> o.w) + (o.y >> o.w) + (o.z >> o.w)) & 1) == colour) {
        const float d = diag[a];
        if (d > 0.0f) v = v + omega * (b[a] - apply_row_from(g, pin, bface, a, o, br, false)) / d;
    }
    pout[a] = v;
}

// weighted Gauss-Seidel half-sweep of one colour (mr_operator.py:193-212),
// colour from the finest-coordinate checkerboard, pin -> pout; one thread
// per cell
__global__ __launch_bounds__(256) void k_msbg_rbgs(MGrid g, const float* __restrict__ pin, float* __restrict__ pout,
                                                   const float* __restrict__ b, const float* __restrict__ bface,
                                                   const float* __restrict__ diag, int colour, float omega) {
    const int a = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    BlockRef br;
    const int4 o = origin_ref(g, a, br);
    rbgs_cell(g, pin, pout, b, bface, diag, colour, omega, a, o, br);
}

// The same half-sweep with thread t on cells 2t and 2t + 1.  When 2t sits at
// an even offset of a block with an even edge, the two are k-neighbours in
// one row and exactly one of them has the active colour: the thread updates
// that one and copies the other, so no lane of a wave idles through the
// coupling walk (with one cell per thread the colour alternates from lane to
// lane and half the lanes wait).  Otherwise both cells take the per-cell
// path, one after the other, which costs latency on grids of small blocks.
// Every cell gets the arithmetic of k_msbg_rbgs.
__global__ __launch_bounds__(256) void k_msbg_rbgs_pairs(MGrid g, const float* __restrict__ pin,
                                                         float* __restrict__ pout, const float* __restrict__ b,
                                                         const float* __restrict__ bface,
                                                         const float* __restrict__ diag, int colour, float omega) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int a0 = 2 * t;
    if (a0 >= g.ncells) return;
    BlockRef br;
    const int4 o0 = origin_ref(g, a0, br);
    if (br.n >= 2 && ((a0 - br.base) & 1) == 0) {
        // a0 + 1 is the next cell of the same row, of the other colour; the
        // pair is read and written as one 8-byte word (the host checks that
        // pin and pout are 8-byte aligned)
        const bool first = (((o0.x >> o0.w) + (o0.y >> o0.w) + (o0.z >> o0.w)) & 1) == colour;
        const int act = first ? a0 : a0 + 1;
        const int4 oa = first ? o0 : make_int4(o0.x, o0.y, o0.z + (1 << o0.w), o0.w);
        float2 v2 = *reinterpret_cast<const float2*>(pin + a0);
        float v = first ? v2.x : v2.y;
        const float d = diag[act];
        if (d > 0.0f) v = v + omega * (b[act] - apply_row_from(g, pin, bface, act, oa, br, false)) / d;
        if (first)
            v2.x = v;
        else
            v2.y = v;
        *reinterpret_cast<float2*>(pout + a0) = v2;
        return;
    }
    rbgs_cell(g, pin, pout, b, bface, diag, colour, omega, a0, o0, br);
    if (a0 + 1 < g.ncells) {
        BlockRef br1;
        const int4 o1 = origin_ref(g, a0 + 1, br1);
        rbgs_cell(g, pin, pout, b, bface, diag, colour, omega, a0 + 1, o1, br1);
    }
}

// All sweeps of the coarsest level in one workgroup: each row's couplings
// (neighbour, T) are gathered into LDS once, then every half-sweep reads LDS
// only.  Same couplings, same order, same arithmetic as k_msbg_rbgs, so the
// result equals the launched half-sweeps bit for bit.  p is read and
// overwritten; the sweep count is even in half-sweeps, so no copy is left.
constexpr int kRowMax = 24;  // 6 faces x a 2x2 patch
constexpr int kCoarsestThreads = 512;

__host__ __device__ inline size_t coarsest_lds_bytes(int n) {
    return (size_t)n * (kRowMax * (sizeof(float) + sizeof(unsigned short)) + 4 * sizeof(float) + 2);
}

__global__ __launch_bounds__(kCoarsestThreads) void k_msbg_coarsest(MGrid g, float* __restrict__ p,
                                                                    const float* __restrict__ b,
                                                                    const float* __restrict__ bface,
                                                                    const float* __restrict__ diag, int sweeps,
                                                                    float omega) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int n = g.ncells;
    float* coef = reinterpret_cast<float*>(lds);
    float* p0 = coef + (size_t)n * kRowMax;
    float* p1 = p0 + n;
    float* bb = p1 + n;
    float* dd = bb + n;
    unsigned short* nbr = reinterpret_cast<unsigned short*>(dd + n);
    unsigned char* cnt = reinterpret_cast<unsigned char*>(nbr + (size_t)n * kRowMax);
    unsigned char* col = cnt + n;
    // entry k of every row is contiguous (k * n + a): a wave's loads of one
    // entry hit consecutive LDS words, free of bank conflicts
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        int k = 0;
        for_each_coupling(g, bface, a, [&](int nb, float t) {
            coef[k * n + a] = t;
            nbr[k * n + a] = (unsigned short)nb;
            ++k;
        });
        cnt[a] = (unsigned char)k;
        const int4 o = origin(g, a);
        col[a] = (unsigned char)(((o.x >> o.w) + (o.y >> o.w) + (o.z >> o.w)) & 1);
        p0[a] = p[a];
        bb[a] = b[a];
        dd[a] = diag[a];
    }
    __syncthreads();
    float* pin = p0;
    float* pout = p1;
    for (int s = 0; s < sweeps; ++s)
        for (int c = 0; c < 2; ++c) {
            for (int a = threadIdx.x; a < n; a += blockDim.x) {
                float v = pin[a];
                if (col[a] == c) {
                    const float d = dd[a];
                    if (d > 0.0f) {
                        const float pa = pin[a];
                        float acc = 0.0f;
                        const int m = cnt[a];
                        for (int k = 0; k < m; ++k) acc += coef[k * n + a] * (pa - pin[nbr[k * n + a]]);
                        v = v + omega * (bb[a] - acc) / d;
                    }
                }
                pout[a] = v;
            }
            __syncthreads();
            float* t = pin;
            pin = pout;
            pout = t;
        }
    for (int a = threadIdx.x; a < n; a += blockDim.x) p[a] = pin[a];
}

// The coarsest level's smoothing as a matrix: `sweeps` red-black sweeps
// from zero followed by mean removal are a linear map b -> p, so column j
// is that map applied to the unit vector e_j.  One workgroup per column,
// the sweeps in FP64 from LDS (couplings gathered as in k_msbg_coarsest);
// op[j * n + t] is entry (t, j).  A V-cycle then applies it with one
// matrix-vector product (k_msbg_coarse_apply) instead of the sweeps.
constexpr int kMsbgCoarseOpMax = 512;

__host__ __device__ inline size_t coarse_op_lds_bytes(int n) {
    return (size_t)n * (kRowMax * (sizeof(float) + sizeof(unsigned short)) + 2 * sizeof(double) + sizeof(float) + 2);
}

__global__ __launch_bounds__(kCoarsestThreads) void k_msbg_coarse_op(MGrid g, const float* __restrict__ bface,
                                                                     const float* __restrict__ diag, int sweeps,
                                                                     double omega, double* __restrict__ op) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int n = g.ncells, j = blockIdx.x;
    double* p0 = reinterpret_cast<double*>(lds);
    double* p1 = p0 + n;
    float* coef = reinterpret_cast<float*>(p1 + n);
    float* dd = coef + (size_t)n * kRowMax;
    unsigned short* nbr = reinterpret_cast<unsigned short*>(dd + n);
    unsigned char* cnt = reinterpret_cast<unsigned char*>(nbr + (size_t)n * kRowMax);
    unsigned char* col = cnt + n;
    __shared__ double sh[16];
    __shared__ double shift;
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
        int k = 0;
        for_each_coupling(g, bface, a, [&](int nb, float t) {
            coef[k * n + a] = t;
            nbr[k * n + a] = (unsigned short)nb;
            ++k;
        });
        cnt[a] = (unsigned char)k;
        const int4 o = origin(g, a);
        col[a] = (unsigned char)(((o.x >> o.w) + (o.y >> o.w) + (o.z >> o.w)) & 1);
        dd[a] = diag[a];
        p0[a] = 0.0;
    }
    __syncthreads();
    double* pin = p0;
    double* pout = p1;
    for (int sw = 0; sw < sweeps; ++sw)
        for (int c = 0; c < 2; ++c) {
            for (int a = threadIdx.x; a < n; a += blockDim.x) {
                double v = pin[a];
                if (col[a] == c && dd[a] > 0.0f) {
                    const double pa = pin[a];
                    double acc = 0.0;
                    const int m = cnt[a];
                    for (int k = 0; k < m; ++k) acc += (double)coef[k * n + a] * (pa - pin[nbr[k * n + a]]);
                    v = pa + omega * ((a == j ? 1.0 : 0.0) - acc) / (double)dd[a];
                }
                pout[a] = v;
            }
            __syncthreads();
            double* t = pin;
            pin = pout;
            pout = t;
        }
    double sm[1] = {0.0};
    for (int a = threadIdx.x; a < n; a += blockDim.x) sm[0] += pin[a];
    block_sum<1>(sm, sh);
    if (threadIdx.x == 0) shift = -(sm[0] / (double)n);
    __syncthreads();
    for (int a = threadIdx.x; a < n; a += blockDim.x) op[(int64_t)j * n + a] = pin[a] + shift;
}

// p = op b in FP64, rounded to float.  A workgroup takes 64 rows; wave w
// sums the columns w, w + 8, w + 16, ... of them (a wave reads 64
// consecutive entries of a column), and the 8 partials of a row are added
// in wave order.
__global__ __launch_bounds__(512) void k_msbg_coarse_apply(const double* __restrict__ op,
                                                           const float* __restrict__ b, float* __restrict__ p,
                                                           int n) {
    __shared__ double bs[kMsbgCoarseOpMax];
    __shared__ double part[8][64];
    for (int j = threadIdx.x; j < n; j += blockDim.x) bs[j] = (double)b[j];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int t = blockIdx.x * 64 + lane;
    double acc = 0.0;
    if (t < n)
        for (int j = w; j < n; j += 8) acc += op[(int64_t)j * n + t] * bs[j];
    part[w][lane] = acc;
    __syncthreads();
    if (w == 0 && t < n) {
        double s = part[0][lane];
        for (int q = 1; q < 8; ++q) s += part[q][lane];
        p[t] = (float)s;
    }
}

// ------------------------------------------------------------------ adaptive relaxation
// Algorithm 2 of the paper as the reference runs it on a GPU
// (adaptive_relax.py:1-241): red-black sweeps restricted to an active block
// set, the per-block effective residual max |delta| diag, a threshold from
// the log-histogram of the first residuals, and re-activation of the blocks
// above it with their face neighbours.  Like k_msbg_rbgs the half-sweep reads
// pin and writes every cell of pout.

// one colour over the active blocks; delta = |update| of the updated cells
// (adaptive_relax.py:84-109)
__global__ __launch_bounds__(256) void k_msbg_rbgs_active(MGrid g, const float* __restrict__ pin,
                                                          float* __restrict__ pout, const float* __restrict__ b,
                                                          const float* __restrict__ bface,
                                                          const float* __restrict__ diag,
                                                          const int* __restrict__ active, int colour, float omega,
                                                          float* __restrict__ delta) {
    const int a = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    float v = pin[a];
    if (active[g.cslot[a]]) {
        const int4 o = origin(g, a);
        if ((((o.x >> o.w) + (o.y >> o.w) + (o.z >> o.w)) & 1) == colour) {
            const float d = diag[a];
            if (d > 0.0f) {
                const float upd = omega * (b[a] - apply_row(g, pin, bface, a, false)) / d;
                v = v + upd;
                delta[a] = fabsf(upd);
            }
        }
    }
    pout[a] = v;
}

// k_msbg_rbgs_active with thread t on cells 2t and 2t + 1, paired as in
// k_msbg_rbgs_pairs (one cell of each colour in one block row, so both share
// the block's active flag); other cells take the per-cell path.  Same
// arithmetic and the same delta writes per cell.
__device__ inline void rbgs_active_cell(const MGrid& g, const float* __restrict__ pin, float* __restrict__ pout,
                                        const float* __restrict__ b, const float* __restrict__ bface,
                                        const float* __restrict__ diag, const int* __restrict__ active,
                                        int colour, float omega, float* __restrict__ delta, int a) {
    float v = pin[a];
    if (active[g.cslot[a]]) {
        BlockRef br;
        const int4 o = origin_ref(g, a, br);
        if ((((o.x >> o.w) + (o.y >> o.w) + (o.z >> o.w)) & 1) == colour) {
            const float d = diag[a];
            if (d > 0.0f) {
                const float upd = omega * (b[a] - apply_row_from(g, pin, bface, a, o, br, false)) / d;
                v = v + upd;
                delta[a] = fabsf(upd);
            }
        }
    }
    pout[a] = v;
}

__global__ __launch_bounds__(256) void k_msbg_rbgs_active_pairs(MGrid g, const float* __restrict__ pin,
                                                                float* __restrict__ pout,
                                                                const float* __restrict__ b,
                                                                const float* __restrict__ bface,
                                                                const float* __restrict__ diag,
                                                                const int* __restrict__ active, int colour,
                                                                float omega, float* __restrict__ delta) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int a0 = 2 * t;
    if (a0 >= g.ncells) return;
    BlockRef br;
    const int4 o0 = origin_ref(g, a0, br);
    if (br.n >= 2 && ((a0 - br.base) & 1) == 0) {
        float2 v2 = *reinterpret_cast<const float2*>(pin + a0);
        if (active[br.slot]) {
            const bool first = (((o0.x >> o0.w) + (o0.y >> o0.w) + (o0.z >> o0.w)) & 1) == colour;
            const int act = first ? a0 : a0 + 1;
            const int4 oa = first ? o0 : make_int4(o0.x, o0.y, o0.z + (1 << o0.w), o0.w);
            const float d = diag[act];
            if (d > 0.0f) {
                const float upd = omega * (b[act] - apply_row_from(g, pin, bface, act, oa, br, false)) / d;
                if (first)
                    v2.x = v2.x + upd;
                else
                    v2.y = v2.y + upd;
                delta[act] = fabsf(upd);
            }
        }
        *reinterpret_cast<float2*>(pout + a0) = v2;
        return;
    }
    rbgs_active_cell(g, pin, pout, b, bface, diag, active, colour, omega, delta, a0);
    if (a0 + 1 < g.ncells) rbgs_active_cell(g, pin, pout, b, bface, diag, active, colour, omega, delta, a0 + 1);
}

// res[slot] = max over the block's cells of delta * diag, 0 when inactive
// (adaptive_relax.py:112-129); one workgroup per block
__global__ __launch_bounds__(256) void k_msbg_block_res(MGrid g, const float* __restrict__ delta,
                                                        const float* __restrict__ diag,
                                                        const int* __restrict__ active, float* __restrict__ res) {
    const int s = blockIdx.x;
    __shared__ float sh[4];
    float m = 0.0f;
    if (active[s]) {
        const int c0 = g.coff[s];
        const int nn = g.n0 >> g.blev[s];
        const int cnt = nn * nn * nn;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) m = fmaxf(m, delta[c0 + i] * diag[c0 + i]);
    }
    m = block_max(m, sh);
    if (threadIdx.x == 0) res[s] = m;
}

// theta = exp(mean + sigma std) of log(res) over the positive entries
// (population std, two passes, FP64 in a fixed order); +inf when none is
// positive, which stops all further activation (adaptive_relax.py:220-231)
__global__ __launch_bounds__(256) void k_msbg_theta(const float* __restrict__ res, int ns, float sigma,
                                                    float* __restrict__ theta) {
    __shared__ double sh[2 * 16];
    double v[2] = {0.0, 0.0};
    for (int s = threadIdx.x; s < ns; s += blockDim.x)
        if (res[s] > 0.0f) {
            v[0] += log((double)res[s]);
            v[1] += 1.0;
        }
    block_sum<2>(v, sh);
    __shared__ double mean_cnt[2];
    if (threadIdx.x == 0) {
        mean_cnt[0] = v[1] > 0.0 ? v[0] / v[1] : 0.0;
        mean_cnt[1] = v[1];
    }
    __syncthreads();
    const double mean = mean_cnt[0], cnt = mean_cnt[1];
    double w[1] = {0.0};
    for (int s = threadIdx.x; s < ns; s += blockDim.x)
        if (res[s] > 0.0f) {
            const double d = log((double)res[s]) - mean;
            w[0] += d * d;
        }
    block_sum<1>(w, sh);
    if (threadIdx.x == 0)
        *theta = cnt > 0.0 ? (float)exp(mean + (double)sigma * sqrt(w[0] / cnt)) : __int_as_float(0x7f800000);
}

// blocks above theta and their face neighbours form the next active set
// (adaptive_relax.py:132-160); the writers all store 1
__global__ __launch_bounds__(256) void k_msbg_update_active(MGrid g, const float* __restrict__ res,
                                                            const float* __restrict__ theta,
                                                            int* __restrict__ active_next) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.nslots) return;
    if (!(res[s] > *theta)) return;
    active_next[s] = 1;
    const int b = g.blk[s];
    const int bz = b % g.nbz, by = (b / g.nbz) % g.nby, bx = b / (g.nbz * g.nby);
    for (int d = 0; d < 6; ++d) {
        const int ax = d >> 1, sg = (d & 1) ? 1 : -1;
        const int nx = bx + (ax == 0 ? sg : 0), ny = by + (ax == 1 ? sg : 0), nz = bz + (ax == 2 ? sg : 0);
        if (nx < 0 || ny < 0 || nz < 0 || nx >= g.nbx || ny >= g.nby || nz >= g.nbz) continue;
        const int ns = g.slot[(nx * g.nby + ny) * g.nbz + nz];
        if (ns >= 0) active_next[ns] = 1;
    }
}

// coarse face = mean of the 2x2 fine faces it covers (mr_multigrid.py:52-89)
__global__ __launch_bounds__(256) void k_msbg_coarsen_beta(MGrid cg, MGrid fg, const float* __restrict__ bfine,
                                                           float* __restrict__ bcoarse) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= cg.ncells) return;
    const int4 o = origin(cg, a);
    const int half = (1 << o.w) >> 1;
    for (int axis = 0; axis < 3; ++axis) {
        const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
        float acc = 0.0f, cnt = 0.0f;
        for (int u = 0; u < 2; ++u)
            for (int v = 0; v < 2; ++v) {
                const int du = u * half, dv = v * half;
                const MCell nc = locate(fg, o.x + comp(a1, 0) * du + comp(a2, 0) * dv,
                                        o.y + comp(a1, 1) * du + comp(a2, 1) * dv,
                                        o.z + comp(a1, 2) * du + comp(a2, 2) * dv);
                if (nc.addr >= 0) {
                    acc += bfine[nc.addr * 3 + axis];
                    cnt += 1.0f;
                }
            }
        bcoarse[a * 3 + axis] = cnt > 0.0f ? acc / cnt : 0.0f;
    }
}

// R = P^T: each coarse cell sums the fine cells it covers, in x-major order:
// eight, or one where the block has the same resolution on both levels (it
// bottomed out at one cell per block) (mr_multigrid.py:92-99)
__global__ __launch_bounds__(256) void k_msbg_restrict(MGrid fg, MGrid cg, const float* __restrict__ rfine,
                                                       float* __restrict__ bcoarse) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= cg.ncells) return;
    const int4 o = origin(cg, a);
    const MCell f0 = locate(fg, o.x, o.y, o.z);
    float acc = 0.0f;
    if (f0.level == o.w) {
        acc = rfine[f0.addr];
    } else {
        const int hf = 1 << f0.level;
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y)
                for (int z = 0; z < 2; ++z)
                    acc += rfine[locate(fg, o.x + x * hf, o.y + y * hf, o.z + z * hf).addr];
    }
    bcoarse[a] = acc;
}

// P: piecewise-constant injection added onto the fine level (mr_multigrid.py:102-109)
__global__ __launch_bounds__(256) void k_msbg_prolong_add(MGrid fg, MGrid cg, const float* __restrict__ pcoarse,
                                                          float* __restrict__ pfine) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= fg.ncells) return;
    const int4 o = origin(fg, a);
    pfine[a] += pcoarse[locate(cg, o.x, o.y, o.z).addr];
}

// cell centre, volume and whether all six neighbours exist
// (msbg.py:144-153, mr_operator.py:215-220)
__global__ __launch_bounds__(256) void k_msbg_geometry(MGrid g, float* __restrict__ centre, float* __restrict__ vol,
                                                       int* __restrict__ interior) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    const int4 o = origin(g, a);
    const int ext = 1 << o.w;
    const float half = (float)ext * 0.5f;
    if (centre) {
        centre[3 * a] = ((float)o.x + half) * g.h0;
        centre[3 * a + 1] = ((float)o.y + half) * g.h0;
        centre[3 * a + 2] = ((float)o.z + half) * g.h0;
    }
    if (vol) {
        const float h = g.h0 * (float)ext;
        vol[a] = h * h * h;
    }
    if (interior) {
        int ok = 1;
        for (int axis = 0; axis < 3; ++axis) {
            const int ex = comp(axis, 0), ey = comp(axis, 1), ez = comp(axis, 2);
            if (locate(g, o.x - ex, o.y - ey, o.z - ez).addr < 0) ok = 0;
            if (locate(g, o.x + ex * ext, o.y + ey * ext, o.z + ez * ext).addr < 0) ok = 0;
        }
        interior[a] = ok;
    }
}

// hits[c] = number of finest coordinates that resolve to cell c
__global__ __launch_bounds__(256) void k_msbg_cover(MGrid g, int* __restrict__ hits) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int fz_n = g.nbz << g.sh0, fy_n = g.nby << g.sh0, fx_n = g.nbx << g.sh0;
    if (t >= (int64_t)fx_n * fy_n * fz_n) return;
    const int fz = (int)(t % fz_n), fy = (int)((t / fz_n) % fy_n), fx = (int)(t / ((int64_t)fz_n * fy_n));
    const MCell c = locate(g, fx, fy, fz);
    if (c.addr >= 0) atomicAdd(hits + c.addr, 1);
}

// FP64 partial sums over blocks, combined in block order by one block
__global__ __launch_bounds__(256) void k_vec_part(int op, const float* __restrict__ x, const float* __restrict__ y,
                                                  int64_t n, double* __restrict__ part) {
    __shared__ double sh[16];
    double v[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (op == 0)
            v[0] += (double)x[i];
        else if (op == 1)
            v[0] += (double)x[i] * (double)y[i];
        else
            v[0] = fmax(v[0], fabs((double)x[i]));
    }
    if (op == 2) {
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        double m = v[0];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if (lane == 0) sh[wid] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = sh[0];
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, sh[w]);
            part[blockIdx.x] = r;
        }
        return;
    }
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// one wave combines the partials: lane t takes t, t + 64, ... in order,
// then a fixed butterfly
__global__ __launch_bounds__(64) void k_vec_comb(int op, const double* __restrict__ part, int np,
                                                 double* __restrict__ out) {
    double r = 0.0;
    for (int b = threadIdx.x; b < np; b += 64) r = op == 2 ? fmax(r, part[b]) : r + part[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(r, o, 64);
        r = op == 2 ? fmax(r, t) : r + t;
    }
    if (threadIdx.x == 0) *out = r;
}

// y = a x + b y (x may be NULL: y = b y + a)
__global__ __launch_bounds__(256) void k_vec_axpby(float a, const float* __restrict__ x, float b, float* __restrict__ y,
                                                   int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = x ? a * x[i] + b * y[i] : b * y[i] + a;
}

// y = a x + b y with a = a_sign * st[ia] and b = st[ib] (or 1 when ib < 0)
// rounded to float, the scalars of the flexible PCG kept on the device; a
// no-op once flag[0] is set
__global__ __launch_bounds__(256) void k_vec_axpby_dev(const double* __restrict__ st, int ia, float a_sign,
                                                       const float* __restrict__ x, int ib,
                                                       float* __restrict__ y, int64_t n,
                                                       const int* __restrict__ flag) {
    if (flag && flag[0]) return;
    const float a = ia >= 0 ? (float)((double)a_sign * st[ia]) : a_sign;
    const float b = ib >= 0 ? (float)st[ib] : 1.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}

// y = 1 y + (float)(-sum / n): the mean removal of mr_multigrid.py:181-184
// with the sum left on the device by pf_vec_reduce
__global__ __launch_bounds__(256) void k_vec_sub_mean(float* __restrict__ y, int64_t n,
                                                      const double* __restrict__ sum) {
    const float a = (float)(-*sum / (double)n);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = 1.0f * y[i] + a;
}

// The scalar steps of the flexible Polak-Ribiere PCG (mr_solver.py:118-165)
// on one thread.  st: 0 rz, 1 alpha, 2 beta, 3 rel, 4 bnorm, 5 tol, 6 qAq,
// 7 a copy of flag[0]; flag[0]: 0 running, 1 stopped at qAq <= 0, 2 converged, 3 stopped at
// rz == 0 or rz_new <= 0.
enum MrcgOp { MRCG_INIT = 0, MRCG_RZ, MRCG_ALPHA, MRCG_REL, MRCG_BETA };

__global__ void k_mrcg_scalar(int op, double* st, int* flag, const double* a, const double* b, double tol) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (op == MRCG_INIT) {  // a = max|b|
        st[4] = *a;
        st[5] = tol;
        st[3] = 1.0;
        flag[0] = 0;
    } else if (op == MRCG_RZ) {
        st[0] = *a;
    } else if (flag[0]) {
        // stopped: the scalars keep their values
    } else if (op == MRCG_ALPHA) {
        st[6] = *a;
        if (*a <= 0.0)
            flag[0] = 1;
        else
            st[1] = st[0] / *a;
    } else if (op == MRCG_REL) {
        st[3] = *a / st[4];
        if (st[3] < st[5]) flag[0] = 2;
    } else if (op == MRCG_BETA) {  // a = <r, z>, b = <r_prev, z>
        const double rz_new = *a;
        if (st[0] == 0.0 || rz_new <= 0.0) {
            flag[0] = 3;
        } else {
            st[2] = fmax((rz_new - *b) / st[0], 0.0);
            st[0] = rz_new;
        }
    }
    st[7] = (double)flag[0];  // the host reads st[0..7] in one copy
}

// z = r / diag where diag > 0, else 0 (mr_solver.py:64-69)
__global__ __launch_bounds__(256) void k_vec_jacobi(const float* __restrict__ r, const float* __restrict__ diag,
                                                    float* __restrict__ z, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float d = diag[i];
        z[i] = d > 0.0f ? r[i] / d : 0.0f;
    }
}

constexpr int kVecBlocks = 1024;

}  // namespace pf

using namespace pf;

static int check_grid(const pf_msbg_grid* g, const char* who) {
    PF_REQUIRE(g, "%s: null grid", who);
    PF_REQUIRE(g->ncells > 0 && g->nslots > 0 && g->n0 == (1 << g->sh0) && g->h0 > 0.f, "%s: bad grid", who);
    PF_REQUIRE(g->level && g->slot && g->blk && g->blev && g->coff && g->cslot, "%s: null grid array", who);
    // face fields are indexed 3 * cell + axis in 32-bit arithmetic
    PF_REQUIRE(g->ncells <= kMsbgMaxCells, "%s: %d cells exceed the limit of %d", who, g->ncells, kMsbgMaxCells);
    return PF_OK;
}

PF_API int pf_msbg_apply(const pf_msbg_grid* g, const float* p, const float* bface, float* out, int32_t want_diag,
                         pf_stream_t stream) {
    int rc = check_grid(g, "pf_msbg_apply");
    if (rc) return rc;
    PF_REQUIRE(bface && out && (want_diag || p), "pf_msbg_apply: null vector");
    k_msbg_apply<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), p, bface, out, want_diag);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_residual(const pf_msbg_grid* g, const float* p, const float* b, const float* bface, float* r,
                            pf_stream_t stream) {
    int rc = check_grid(g, "pf_msbg_residual");
    if (rc) return rc;
    PF_REQUIRE(p && b && bface && r, "pf_msbg_residual: null vector");
    k_msbg_residual<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), p, b, bface, r);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_rbgs(const pf_msbg_grid* g, const float* pin, float* pout, const float* b, const float* bface,
                        const float* diag, int32_t colour, float omega, pf_stream_t stream) {
    int rc = check_grid(g, "pf_msbg_rbgs");
    if (rc) return rc;
    PF_REQUIRE(pin && pout && b && bface && diag, "pf_msbg_rbgs: null vector");
    PF_REQUIRE(pin != pout, "pf_msbg_rbgs: pin and pout must be distinct");
    PF_REQUIRE(colour == 0 || colour == 1, "pf_msbg_rbgs: colour must be 0 or 1");
    if (g->ncells >= kMsbgPairMinCells && ((reinterpret_cast<uintptr_t>(pin) | reinterpret_cast<uintptr_t>(pout)) & 7) == 0)
        k_msbg_rbgs_pairs<<<cdiv(cdiv(g->ncells, 2), 256), 256, 0, S(stream)>>>(mgrid(*g), pin, pout, b, bface, diag,
                                                                               colour, omega);
    else
        k_msbg_rbgs<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), pin, pout, b, bface, diag, colour, omega);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_rbgs_active(const pf_msbg_grid* g, const float* pin, float* pout, const float* b,
                               const float* bface, const float* diag, const int32_t* active, int32_t colour,
                               float omega, float* delta, pf_stream_t stream) {
    int rc = check_grid(g, "pf_msbg_rbgs_active");
    if (rc) return rc;
    PF_REQUIRE(pin && pout && b && bface && diag && active && delta, "pf_msbg_rbgs_active: null vector");
    PF_REQUIRE(pin != pout, "pf_msbg_rbgs_active: pin and pout must be distinct");
    PF_REQUIRE(colour == 0 || colour == 1, "pf_msbg_rbgs_active: colour must be 0 or 1");
    if (g->ncells >= kMsbgPairMinCells && ((reinterpret_cast<uintptr_t>(pin) | reinterpret_cast<uintptr_t>(pout)) & 7) == 0)
        k_msbg_rbgs_active_pairs<<<cdiv(cdiv(g->ncells, 2), 256), 256, 0, S(stream)>>>(
            mgrid(*g), pin, pout, b, bface, diag, active, colour, omega, delta);
    else
        k_msbg_rbgs_active<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), pin, pout, b, bface, diag, active,
                                                                        colour, omega, delta);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_relax_update(const pf_msbg_grid* g, const float* delta, const float* diag, const int32_t* active,
                                float* res, float* theta, int32_t set_theta, float theta_sigma, int32_t* active_next,
                                pf_stream_t stream) {
    int rc = check_grid(g, "pf_msbg_relax_update");
    if (rc) return rc;
    PF_REQUIRE(delta && diag && active && res && theta && active_next, "pf_msbg_relax_update: null pointer");
    PF_REQUIRE(active_next != active, "pf_msbg_relax_update: active_next must be a separate buffer");
    k_msbg_block_res<<<g->nslots, 256, 0, S(stream)>>>(mgrid(*g), delta, diag, active, res);
    PF_CHECK_LAUNCH();
    if (set_theta) {
        k_msbg_theta<<<1, 256, 0, S(stream)>>>(res, g->nslots, theta_sigma, theta);
        PF_CHECK_LAUNCH();
    }
    rc = fill_u32(active_next, g->nslots, 0u, S(stream));
    if (rc) return rc;
    k_msbg_update_active<<<cdiv(g->nslots, 256), 256, 0, S(stream)>>>(mgrid(*g), res, theta, active_next);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int32_t pf_msbg_coarsest_max_cells(void) {
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    int n = (int)(lim / coarsest_lds_bytes(1));
    return n < 65536 ? n : 65535;  // neighbour indices are 16-bit
}

PF_API int pf_msbg_coarsest_smooth(const pf_msbg_grid* g, float* p, const float* b, const float* bface,
                                   const float* diag, int32_t sweeps, float omega, pf_stream_t stream) {
    int rc = check_grid(g, "pf_msbg_coarsest_smooth");
    if (rc) return rc;
    PF_REQUIRE(p && b && bface && diag && sweeps >= 0, "pf_msbg_coarsest_smooth: bad arguments");
    PF_REQUIRE(g->ncells <= pf_msbg_coarsest_max_cells(),
               "pf_msbg_coarsest_smooth: %d cells exceed the workgroup's LDS", g->ncells);
    const size_t bytes = coarsest_lds_bytes(g->ncells);
    if (bytes > 65536) {
        const hipError_t e = hipFuncSetAttribute((const void*)k_msbg_coarsest,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        PF_REQUIRE(e == hipSuccess, "pf_msbg_coarsest_smooth: cannot raise the LDS limit to %zu bytes", bytes);
    }
    k_msbg_coarsest<<<1, kCoarsestThreads, bytes, S(stream)>>>(mgrid(*g), p, b, bface, diag, sweeps, omega);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int32_t pf_msbg_coarse_op_max_cells(void) {
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    const int n = (int)(lim / coarse_op_lds_bytes(1));
    return n < kMsbgCoarseOpMax ? n : kMsbgCoarseOpMax;
}

PF_API int pf_msbg_coarse_op(const pf_msbg_grid* g, const float* bface, const float* diag, int32_t sweeps,
                             float omega, double* op, pf_stream_t stream) {
    int rc = check_grid(g, "pf_msbg_coarse_op");
    if (rc) return rc;
    PF_REQUIRE(bface && diag && op && sweeps >= 0, "pf_msbg_coarse_op: bad arguments");
    PF_REQUIRE(g->ncells <= pf_msbg_coarse_op_max_cells(), "pf_msbg_coarse_op: %d cells exceed the limit",
               g->ncells);
    const size_t bytes = coarse_op_lds_bytes(g->ncells);
    if (bytes > 65536) {
        const hipError_t e = hipFuncSetAttribute((const void*)k_msbg_coarse_op,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        PF_REQUIRE(e == hipSuccess, "pf_msbg_coarse_op: cannot raise the LDS limit to %zu bytes", bytes);
    }
    k_msbg_coarse_op<<<g->ncells, kCoarsestThreads, bytes, S(stream)>>>(mgrid(*g), bface, diag, sweeps,
                                                                         (double)omega, op);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_coarse_apply(const double* op, const float* b, float* p, int32_t n, pf_stream_t stream) {
    PF_REQUIRE(op && b && p && n > 0 && n <= kMsbgCoarseOpMax, "pf_msbg_coarse_apply: bad arguments");
    k_msbg_coarse_apply<<<cdiv(n, 64), 512, 0, S(stream)>>>(op, b, p, n);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_coarsen_beta(const pf_msbg_grid* cg, const pf_msbg_grid* fg, const float* bfine, float* bcoarse,
                                pf_stream_t stream) {
    int rc = check_grid(cg, "pf_msbg_coarsen_beta");
    if (rc) return rc;
    rc = check_grid(fg, "pf_msbg_coarsen_beta");
    if (rc) return rc;
    PF_REQUIRE(bfine && bcoarse, "pf_msbg_coarsen_beta: null vector");
    k_msbg_coarsen_beta<<<cdiv(cg->ncells, 256), 256, 0, S(stream)>>>(mgrid(*cg), mgrid(*fg), bfine, bcoarse);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_restrict(const pf_msbg_grid* fg, const pf_msbg_grid* cg, const float* rfine, float* bcoarse,
                            pf_stream_t stream) {
    int rc = check_grid(fg, "pf_msbg_restrict");
    if (rc) return rc;
    rc = check_grid(cg, "pf_msbg_restrict");
    if (rc) return rc;
    PF_REQUIRE(rfine && bcoarse, "pf_msbg_restrict: null vector");
    k_msbg_restrict<<<cdiv(cg->ncells, 256), 256, 0, S(stream)>>>(mgrid(*fg), mgrid(*cg), rfine, bcoarse);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_prolong_add(const pf_msbg_grid* fg, const pf_msbg_grid* cg, const float* pcoarse, float* pfine,
                               pf_stream_t stream) {
    int rc = check_grid(fg, "pf_msbg_prolong_add");
    if (rc) return rc;
    rc = check_grid(cg, "pf_msbg_prolong_add");
    if (rc) return rc;
    PF_REQUIRE(pcoarse && pfine, "pf_msbg_prolong_add: null vector");
    k_msbg_prolong_add<<<cdiv(fg->ncells, 256), 256, 0, S(stream)>>>(mgrid(*fg), mgrid(*cg), pcoarse, pfine);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_geometry(const pf_msbg_grid* g, float* centre, float* volume, int32_t* interior,
                            pf_stream_t stream) {
    int rc = check_grid(g, "pf_msbg_geometry");
    if (rc) return rc;
    k_msbg_geometry<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), centre, volume, interior);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_msbg_cover(const pf_msbg_grid* g, int32_t* hits, pf_stream_t stream) {
    int rc = check_grid(g, "pf_msbg_cover");
    if (rc) return rc;
    PF_REQUIRE(hits, "pf_msbg_cover: null hits");
    const int64_t nf = ((int64_t)g->nbx << g->sh0) * ((int64_t)g->nby << g->sh0) * ((int64_t)g->nbz << g->sh0);
    PF_REQUIRE(nf < ((int64_t)1 << 31), "pf_msbg_cover: too many finest cells");
    rc = fill_u32(hits, g->ncells, 0u, S(stream));
    if (rc) return rc;
    k_msbg_cover<<<cdiv(nf, 256), 256, 0, S(stream)>>>(mgrid(*g), hits);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API size_t pf_vec_reduce_bytes(void) { return (size_t)(kVecBlocks + 1) * sizeof(double); }

PF_API int pf_vec_reduce(int32_t op, const float* x, const float* y, int64_t n, double* work, pf_stream_t stream) {
    PF_REQUIRE(op >= 0 && op <= 2, "pf_vec_reduce: op must be 0 (sum), 1 (dot) or 2 (max |x|)");
    PF_REQUIRE(x && work && (op != 1 || y), "pf_vec_reduce: null vector");
    PF_REQUIRE(n > 0, "pf_vec_reduce: empty vector");
    const int64_t nb64 = cdiv(n, 256 * 4);
    const int nb = (int)(nb64 < kVecBlocks ? nb64 : kVecBlocks);
    k_vec_part<<<nb, 256, 0, S(stream)>>>(op, x, y, n, work + 1);
    PF_CHECK_LAUNCH();
    k_vec_comb<<<1, 64, 0, S(stream)>>>(op, work + 1, nb, work);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_vec_axpby(float a, const float* x, float b, float* y, int64_t n, pf_stream_t stream) {
    PF_REQUIRE(y && n > 0, "pf_vec_axpby: null or empty y");
    const int64_t nb64 = cdiv(n, 256 * 4);
    const int nb = (int)(nb64 < kVecBlocks ? nb64 : kVecBlocks);
    k_vec_axpby<<<nb, 256, 0, S(stream)>>>(a, x, b, y, n);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_vec_axpby_dev(const double* st, int32_t ia, float a_sign, const float* x, int32_t ib, float* y,
                            int64_t n, const int32_t* flag, pf_stream_t stream) {
    PF_REQUIRE(st && x && y && n > 0, "pf_vec_axpby_dev: null or empty vector");
    PF_REQUIRE(ia < 8 && ib < 8, "pf_vec_axpby_dev: scalar slot out of range");
    const int64_t nb64 = cdiv(n, 256 * 4);
    const int nb = (int)(nb64 < kVecBlocks ? nb64 : kVecBlocks);
    k_vec_axpby_dev<<<nb, 256, 0, S(stream)>>>(st, ia, a_sign, x, ib, y, n, flag);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_vec_remove_mean(float* x, int64_t n, double* work, pf_stream_t stream) {
    int rc = pf_vec_reduce(0, x, nullptr, n, work, stream);
    if (rc) return rc;
    const int64_t nb64 = cdiv(n, 256 * 4);
    const int nb = (int)(nb64 < kVecBlocks ? nb64 : kVecBlocks);
    k_vec_sub_mean<<<nb, 256, 0, S(stream)>>>(x, n, work);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mrcg_scalar(int32_t op, double* st, int32_t* flag, const double* a, const double* b, double tol,
                          pf_stream_t stream) {
    PF_REQUIRE(st && flag && (op < MRCG_ALPHA || a), "pf_mrcg_scalar: null pointer");
    PF_REQUIRE(op >= MRCG_INIT && op <= MRCG_BETA && (op != MRCG_BETA || b) && (op > MRCG_RZ || a),
               "pf_mrcg_scalar: bad op");
    k_mrcg_scalar<<<1, 64, 0, S(stream)>>>(op, st, flag, a, b, tol);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_vec_jacobi(const float* r, const float* diag, float* z, int64_t n, pf_stream_t stream) {
    PF_REQUIRE(r && diag && z && n > 0, "pf_vec_jacobi: null or empty vector");
    const int64_t nb64 = cdiv(n, 256 * 4);
    const int nb = (int)(nb64 < kVecBlocks ? nb64 : kVecBlocks);
    k_vec_jacobi<<<nb, 256, 0, S(stream)>>>(r, diag, z, n);
    PF_CHECK_LAUNCH();
    return PF_OK;
}
