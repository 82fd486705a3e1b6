// Particle-side kernels: binning + stable counting sort, tiled P2G with LDS
// accumulators, fused G2P + RK3 advection + wall reflection (+ solids), and
// the CFL speed reduction.
#include <mutex>

#include "pf_common.hpp"

namespace pf {

int fill_u32(void* dst, int64_t n, uint32_t v, hipStream_t s);

constexpr int T = PF_TILE;
constexpr int T3 = T * T * T;

struct TileGeom {
    int nx, ny, nz, ntx, nty, ntz;
};

static TileGeom tile_geom(int nx, int ny, int nz) {
    TileGeom g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.ntx = (nx + T - 1) / T;
    g.nty = (ny + T - 1) / T;
    g.ntz = (nz + T - 1) / T;
    return g;
}

static int64_t n_tiles(const TileGeom& g) { return (int64_t)g.ntx * g.nty * g.ntz; }

// ============================================================== binning
// Particles are visited in `order` (the previous step's sorted permutation,
// or index order).  Particles move at most about one cell per step, so equal
// keys arrive in runs of consecutive lanes: each run takes its slots with one
// atomic of its length, and ranks inside the run follow lane order.  Inactive
// particles (pstate != 0) go to the extra bucket nkeys, after all cells.
__device__ inline uint32_t cell_key(float x, float y, float z, float dx, const TileGeom& g) {
    int i = (int)floorf(x / dx), j = (int)floorf(y / dx), k = (int)floorf(z / dx);
    i = min(max(i, 0), g.nx - 1);
    j = min(max(j, 0), g.ny - 1);
    k = min(max(k, 0), g.nz - 1);
    const int tile = ((i / T) * g.nty + (j / T)) * g.ntz + (k / T);
    const int loc = ((i % T) * T + (j % T)) * T + (k % T);
    return (uint32_t)tile * T3 + (uint32_t)loc;
}

__global__ __launch_bounds__(256) void k_bin(const float* __restrict__ px,
                                             const int* __restrict__ pstate, int64_t n,
                                             const uint32_t* __restrict__ order, TileGeom g,
                                             float dx, uint32_t nkeys, uint32_t* __restrict__ keys,
                                             uint32_t* __restrict__ rank, int* __restrict__ count) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = idx < n;
    uint32_t bucket = 0xFFFFFFFEu;  // lanes past the end form their own run
    bool active = false;
    if (valid) {
        const uint32_t p = order ? order[idx] : (uint32_t)idx;
        // the state and the position in one round trip (the position is read
        // whatever the state)
        const int st = pstate[p];
        const float x = px[3 * (int64_t)p], y = px[3 * (int64_t)p + 1], z = px[3 * (int64_t)p + 2];
        asm volatile("" ::"v"(st), "v"(x), "v"(y), "v"(z));
        if (st == 0) {
            active = true;
            bucket = cell_key(x, y, z, dx, g);
        } else {
            bucket = nkeys;
        }
        keys[idx] = active ? bucket : 0xFFFFFFFFu;
    }
    // runs of equal buckets among consecutive lanes
    const uint32_t prev = __shfl_up(bucket, 1, 64);
    const bool head = lane == 0 || prev != bucket;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
    const int run0 = 63 - __clzll(heads & upto);
    const unsigned long long after = heads & ~upto;
    const int run1 = after ? (__ffsll((long long)after) - 1) : 64;
    int base = 0;
    if (valid && lane == run0) base = atomicAdd(&count[bucket], run1 - run0);
    base = __shfl(base, run0, 64);
    if (valid) rank[idx] = (uint32_t)(base + lane - run0);
}

__global__ void k_copy_count(const int* __restrict__ start, int64_t nkeys, int* n_active) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *n_active = start[nkeys];
}

// ============================================================== exclusive scan
constexpr int SCAN_THREADS = 256, SCAN_ITEMS = 16, SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const int* __restrict__ in, int64_t n,
                                                              int* __restrict__ block_sums) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        const int64_t i = base + q * SCAN_THREADS + threadIdx.x;
        if (i < n) s += in[i];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ int sh[SCAN_THREADS / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < SCAN_THREADS / 64; ++w) t += sh[w];
        block_sums[blockIdx.x] = t;
    }
}

// Exclusive scan of block sums in one workgroup; offsets[nb] = total.
__global__ __launch_bounds__(1024) void k_scan_blocks(const int* __restrict__ block_sums, int nb,
                                                      int* __restrict__ offsets) {
    __shared__ int sh[1024];
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 1024) {
        const int i = c0 + threadIdx.x;
        const int v = i < nb ? block_sums[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) offsets[i] = carry + sh[threadIdx.x] - v;
        const int total = sh[1023];
        __syncthreads();
        carry += total;
    }
    if (threadIdx.x == 0) offsets[nb] = carry;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int* __restrict__ in, int64_t n,
                                                             const int* __restrict__ offsets, int nb,
                                                             int* __restrict__ out) {
    __shared__ int tile[SCAN_TILE];
    __shared__ int wsum[SCAN_THREADS / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        const int e = q * SCAN_THREADS + threadIdx.x;
        tile[e] = (base + e < n) ? in[base + e] : 0;
    }
    __syncthreads();
    int loc[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        loc[q] = s;
        s += tile[threadIdx.x * SCAN_ITEMS + q];
    }
    // inclusive wave scan of per-thread totals
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int wbase = 0;
    for (int w = 0; w < wid; ++w) wbase += wsum[w];
    const int excl = offsets[blockIdx.x] + wbase + inc - s;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) tile[threadIdx.x * SCAN_ITEMS + q] = excl + loc[q];
    __syncthreads();
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        const int e = q * SCAN_THREADS + threadIdx.x;
        if (base + e < n) out[base + e] = tile[e];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = offsets[nb];
}

int64_t scan_tmp_ints(int64_t n) { return 2 * (int64_t)cdiv(n, SCAN_TILE) + 1; }

// out[0..n] = exclusive prefix sum of in[0..n), out[n] = total; tmp holds
// scan_tmp_ints(n) ints
int exclusive_scan(const int* in, int64_t n, int* out, int* tmp, hipStream_t s) {
    const int nb = (int)cdiv(n, SCAN_TILE);
    int* block_sums = tmp;
    int* offsets = tmp + nb;
    k_scan_reduce<<<nb, SCAN_THREADS, 0, s>>>(in, n, block_sums);
    PF_CHECK_LAUNCH();
    k_scan_blocks<<<1, 1024, 0, s>>>(block_sums, nb, offsets);
    PF_CHECK_LAUNCH();
    k_scan_apply<<<nb, SCAN_THREADS, 0, s>>>(in, n, offsets, nb, out);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

// ============================================================== scatter + per-key order
__global__ __launch_bounds__(256) void k_scatter(const uint32_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ rank, int64_t n,
                                                 const uint32_t* __restrict__ order, uint32_t nkeys,
                                                 const int* __restrict__ start,
                                                 uint32_t* __restrict__ perm) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const uint32_t key = keys[idx];
    const uint32_t bucket = key == 0xFFFFFFFFu ? nkeys : key;
    perm[start[bucket] + rank[idx]] = order ? order[idx] : (uint32_t)idx;
}

// Slots inside one key were handed out by atomics; sorting each key's slice
// by particle index makes the permutation unique (a stable sort by key).
// One wave per slab of T*T keys (a tile's x-slab, contiguous in `perm`):
// the slab is staged in LDS, each lane insertion-sorts one key's slice there,
// and the slab is written back.  Slabs larger than the LDS buffer
// (pathological clustering) are sorted in place in global memory.  A key
// holding more than kMaxSortLen particles keeps its slot order, so the cost
// stays bounded; P2G does not depend on the order within a key.
constexpr int SLAB_KEYS = T * T;
constexpr int SLAB_CAP = 4096;
constexpr int kMaxSortLen = 1024;

__device__ inline void insertion_sort(uint32_t* v, int n) {
    if (n > kMaxSortLen) return;
    for (int a = 1; a < n; ++a) {
        const uint32_t x = v[a];
        int b = a - 1;
        while (b >= 0 && v[b] > x) {
            v[b + 1] = v[b];
            --b;
        }
        v[b + 1] = x;
    }
}

__global__ __launch_bounds__(64) void k_key_sort(const int* __restrict__ start,
                                                 uint32_t* __restrict__ perm) {
    __shared__ uint32_t buf[SLAB_CAP];
    const int key0 = blockIdx.x * SLAB_KEYS;
    const int s0 = start[key0], e0 = start[key0 + SLAB_KEYS];
    const int n = e0 - s0;
    const int lane = threadIdx.x;
    const int ks = start[key0 + lane] - s0, ke = start[key0 + lane + 1] - s0;
    if (n <= SLAB_CAP) {
        for (int e = lane; e < n; e += 64) buf[e] = perm[s0 + e];
        __syncthreads();
        insertion_sort(buf + ks, ke - ks);
        __syncthreads();
        for (int e = lane; e < n; e += 64) perm[s0 + e] = buf[e];
    } else {
        insertion_sort(perm + s0 + ks, ke - ks);
    }
}

// ============================================================== P2G
// Tile kernel: one workgroup per T^3-cell tile accumulates its particles'
// eq. (6) contributions into LDS over the tile plus a one-cell halo, then
// stores that region to a staging buffer with plain coalesced stores.  A
// gather kernel sums, for every face and cell, the <= 8 tile regions that
// overlap it, in a fixed tile order: no global atomics and no pre-zeroing of
// the outputs, and the cross-tile sums are reproducible.
struct P2GArgs {
    const float* px;
    const float* pv;
    const float* pmass;
    const uint32_t* perm;
    const int* start;
    TileGeom g;
    float dx, inv_r2;
    unsigned long long* stage;
    const float* bounds;  // [max|v|, max m, max m|v|_inf]
    const int* cmax;      // max particles in one key
};

// Fixed-point accumulation.  gfx950 executes LDS float atomics (ds_add_f32)
// far slower than integer ones, so every contribution q is added as the
// 64-bit integer round(q * 2^E) with ds_add_u64.  One exponent per quantity
// (mass, momentum, weight) for the whole grid, chosen from device-side bounds
// so that one contribution stays below 2^50 and a face's full sum (at most
// 27 * cmax contributors) below 2^62: no overflow, resolution 2^-E.  Integer
// sums do not depend on the order of the adds, so P2G is bit-reproducible.
struct FixScale {
    float sm, su, sw;        // 2^E for mass, momentum, weight
    float ru;                // su / sm
    double im, iu, iw;       // 2^-E
};

__host__ __device__ inline int fix_exponent(double per_item, double n_items) {
    // largest E with per_item * 2^E <= 2^50 and n_items * per_item * 2^E <= 2^62
    if (!(per_item > 0.0) || !(per_item < 1e300)) return 0;
    int e1, e2;
    frexp(per_item, &e1);                          // per_item < 2^e1
    frexp(per_item * fmax(n_items, 1.0), &e2);     // total < 2^e2
    int e = 50 - e1;
    if (62 - e2 < e) e = 62 - e2;
    return e < -100 ? -100 : (e > 100 ? 100 : e);
}

__device__ inline FixScale fix_scale(const float* bounds, const int* cmax) {
    const double nsup = 27.0 * (double)max(*cmax, 1);
    int em = fix_exponent((double)bounds[1], nsup);
    int eu = fix_exponent((double)bounds[2], nsup);
    // keep su / sm a normal float (a lower exponent only costs precision)
    if (eu > em + 126) eu = em + 126;
    if (em > eu + 126) em = eu + 126;
    const int ew = fix_exponent(1.0, nsup);
    FixScale f;
    f.sm = ldexpf(1.0f, em);
    f.su = ldexpf(1.0f, eu);
    f.sw = ldexpf(1.0f, ew);
    f.ru = ldexpf(1.0f, eu - em);
    f.im = ldexp(1.0, -em);
    f.iu = ldexp(1.0, -eu);
    f.iw = ldexp(1.0, -ew);
    return f;
}

// round(q) as a two's-complement 64-bit integer, |q| < 2^51 (1.5 * 2^52
// aligns the binary point of a double with its last mantissa bit)
__device__ inline unsigned long long fix_round(float q) {
    const double d = (double)q + 6755399441055744.0;
    return (unsigned long long)(__double_as_longlong(d) - 0x4338000000000000LL);
}
// round(q * s)
__device__ inline unsigned long long to_fix(float q, float s) { return fix_round(q * s); }

constexpr int E = T + 2;  // tangential / cell extent of a tile region (1-cell halo)
constexpr int F = T + 1;  // normal extent of a face region
constexpr int NXR = F * E * E, NYR = E * F * E, NZR = E * E * F, NCR = E * E * E;
// LDS / staging layout of one tile region
constexpr int O_XM = 0, O_XU = NXR, O_YM = 2 * NXR, O_YU = 2 * NXR + NYR;
constexpr int O_ZM = 2 * NXR + 2 * NYR, O_ZU = O_ZM + NZR, O_CM = O_ZU + NZR, O_CW = O_CM + NCR;
constexpr int REG = O_CM + NCR;        // floats per tile without the weight sum
constexpr int REG_W = O_CW + NCR;      // with it

__device__ inline float cube_w(float d2, float inv_r2) {
    float t = fmaxf(1.0f - d2 * inv_r2, 0.0f);
    return t * t * t;
}

// Per-axis data of one particle: squared distances to the two normal
// (o = 0) and the two tangential (o = 1/2) candidate planes, and the
// region-local index of the first candidate.  A candidate outside the tile
// region (only possible for particles outside the domain) gets a squared
// distance that makes its weight zero.
struct AxisCand {
    float n0, n1, t0, t1;
    int ln, lt;
};

__device__ inline AxisCand axis_cand(float x, float dx, int origin) {
    AxisCand c;
    const float gx = x / dx;
    const int in = (int)floorf(gx);
    const int it = (int)floorf(gx - 0.5f);
    const float fn0 = (float)in * dx - x;
    const float fn1 = (float)(in + 1) * dx - x;
    const float ft0 = ((float)it + 0.5f) * dx - x;
    const float ft1 = ((float)(it + 1) + 0.5f) * dx - x;
    c.ln = in - origin;
    c.lt = it - (origin - 1);
    const float big = 1e30f;
    c.n0 = ((unsigned)c.ln < (unsigned)F) ? fn0 * fn0 : big;
    c.n1 = ((unsigned)(c.ln + 1) < (unsigned)F) ? fn1 * fn1 : big;
    c.t0 = ((unsigned)c.lt < (unsigned)E) ? ft0 * ft0 : big;
    c.t1 = ((unsigned)(c.lt + 1) < (unsigned)E) ? ft1 * ft1 : big;
    return c;
}

// Accumulate one particle into the 2x2x2 candidates of one staggered grid.
// q*[2] are per-axis squared distances; base is the LDS offset of candidate
// (0,0,0); SX, SY are the LDS strides of the first / second index.
// ms = m * sm and us = u * su / sm fold the power-of-two fixed-point scales
// into the particle: scaling by 2^E commutes with rounding, so w * ms is
// round(w m) 2^Em and (w * ms) * us is round(round(w m) u) 2^Eu, the bits
// of scaling each rounded product, with two multiplies fewer.
template <int SX, int SY, bool MOM>
__device__ inline void splat8(const float (&qx)[2], const float (&qy)[2], const float (&qz)[2],
                              int base, float ms, float us, float inv_r2,
                              unsigned long long* sm, unsigned long long* su) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const float pxy = qx[a] + qy[b];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const float w = cube_w(pxy + qz[c], inv_r2);
                if (w > 0.0f) {
                    const int addr = base + a * SX + b * SY + c;
                    const float wm = w * ms;
                    atomicAdd(&sm[addr], fix_round(wm));
                    if (MOM) atomicAdd(&su[addr], fix_round(wm * us));
                }
            }
        }
}

struct PData {
    float x, y, z, vx, vy, vz, m;
};

__device__ inline PData load_particle(const P2GArgs& a, uint32_t p) {
    PData d;
    const int64_t b = 3 * (int64_t)p;
    d.x = a.px[b];
    d.y = a.px[b + 1];
    d.z = a.px[b + 2];
    d.vx = a.pv[b];
    d.vy = a.pv[b + 1];
    d.vz = a.pv[b + 2];
    d.m = a.pmass[p];
    return d;
}

template <bool WSUM>
__device__ inline void splat_particle(const PData& d, float dx, float inv_r2, int X0, int Y0, int Z0,
                                      const FixScale& fs, unsigned long long* s) {
    const AxisCand cx = axis_cand(d.x, dx, X0), cy = axis_cand(d.y, dx, Y0), cz = axis_cand(d.z, dx, Z0);
    const float nx_[2] = {cx.n0, cx.n1}, tx_[2] = {cx.t0, cx.t1};
    const float ny_[2] = {cy.n0, cy.n1}, ty_[2] = {cy.t0, cy.t1};
    const float nz_[2] = {cz.n0, cz.n1}, tz_[2] = {cz.t0, cz.t1};
    const float ms = d.m * fs.sm;
    // x-faces: region [F][E][E], origin (X0, Y0-1, Z0-1)
    splat8<E * E, E, true>(nx_, ty_, tz_, (cx.ln * E + cy.lt) * E + cz.lt, ms, d.vx * fs.ru, inv_r2,
                           s + O_XM, s + O_XU);
    // y-faces: region [E][F][E], origin (X0-1, Y0, Z0-1)
    splat8<F * E, E, true>(tx_, ny_, tz_, (cx.lt * F + cy.ln) * E + cz.lt, ms, d.vy * fs.ru, inv_r2,
                           s + O_YM, s + O_YU);
    // z-faces: region [E][E][F], origin (X0-1, Y0-1, Z0)
    splat8<E * F, F, true>(tx_, ty_, nz_, (cx.lt * E + cy.lt) * F + cz.ln, ms, d.vz * fs.ru, inv_r2,
                           s + O_ZM, s + O_ZU);
    // cells: region [E][E][E], origin (X0-1, Y0-1, Z0-1)
    if (WSUM) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const float pxy = tx_[i] + ty_[j];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const float w = cube_w(pxy + tz_[k], inv_r2);
                    if (w > 0.0f) {
                        const int addr = ((cx.lt + i) * E + (cy.lt + j)) * E + cz.lt + k;
                        atomicAdd(&s[O_CM + addr], fix_round(w * ms));
                        atomicAdd(&s[O_CW + addr], to_fix(w, fs.sw));
                    }
                }
            }
    } else {
        splat8<E * E, E, false>(tx_, ty_, tz_, (cx.lt * E + cy.lt) * E + cz.lt, ms, 0.f, inv_r2,
                                s + O_CM, nullptr);
    }
}

// Two cells of the tile per lane.  The LDS region (51-59 KB) allows three
// workgroups per CU; one lane per cell (512 lanes, 6 waves per SIMD instead
// of 3) measured 18 % slower: more waves of one tile region put more
// same-address LDS atomics in flight at once.
constexpr int kP2GThreads = 256;

template <bool WSUM>
__global__ __launch_bounds__(kP2GThreads) void k_p2g(P2GArgs a) {
    constexpr int R = WSUM ? REG_W : REG;
    __shared__ unsigned long long s[R];

    const TileGeom g = a.g;
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int tz = tile % g.ntz, ty = (tile / g.ntz) % g.nty, tx = tile / (g.ntz * g.nty);
    const int X0 = tx * T, Y0 = ty * T, Z0 = tz * T;
    const FixScale fs = fix_scale(a.bounds, a.cmax);

    for (int e = threadIdx.x; e < R; e += blockDim.x) s[e] = 0ull;
    __syncthreads();

    const float dx = a.dx, inv_r2 = a.inv_r2;
    const int key0 = tile * T3;
    // lanes of a wave take different cells of one x-slab
    for (int lc = threadIdx.x; lc < T3; lc += blockDim.x) {
        const int s0 = a.start[key0 + lc], e0 = a.start[key0 + lc + 1];
        if (s0 >= e0) continue;
        // software pipeline: the next particle's data and the index after it
        // are in flight while the current particle is splatted
        uint32_t p_after = (s0 + 1 < e0) ? a.perm[s0 + 1] : 0u;
        PData nxt = load_particle(a, a.perm[s0]);
        for (int q = s0; q < e0; ++q) {
            const PData cur = nxt;
            if (q + 1 < e0) {
                nxt = load_particle(a, p_after);
                if (q + 2 < e0) p_after = a.perm[q + 2];
            }
            splat_particle<WSUM>(cur, dx, inv_r2, X0, Y0, Z0, fs, s);
        }
    }
    __syncthreads();
    unsigned long long* out = a.stage + (int64_t)tile * R;
    for (int e = threadIdx.x; e < R; e += blockDim.x) out[e] = s[e];
}

// max particles in one key (bounds the number of contributors to one sum).
// One atomic per block: one per wave put 4096 same-address atomics into a
// 128^3 launch, which then took 51 us for 8 MB of reads.
__global__ __launch_bounds__(256) void k_max_count(const int* __restrict__ start, int64_t nkeys,
                                                   int* cmax) {
    __shared__ int sh[4];
    int m = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nkeys;
         k += (int64_t)gridDim.x * blockDim.x)
        m = max(m, start[k + 1] - start[k]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(cmax, max(max(sh[0], sh[1]), max(sh[2], sh[3])));
}

// Tiles whose region holds index i of an axis.  A normal (face-normal) axis
// of tile t spans [tT, tT+T], a tangential / cell axis [tT-1, tT+T], so an
// index lies in one or two tiles.  Candidate 0 always exists; candidate 1
// exists when `two`, and otherwise repeats candidate 0 (an address that is
// valid to read; its value is masked out of the sums).
struct AxisTiles {
    int t0, l0, t1, l1;  // tile index and region-local index
    bool two;
};

__device__ inline AxisTiles axis_normal(int i, int nt) {
    const int q = i / T, r = i - q * T;
    AxisTiles a;
    a.t0 = q < nt ? q : q - 1;  // q == nt only for the last face, r == 0
    a.l0 = q < nt ? r : r + T;
    a.two = r == 0 && q >= 1 && q < nt;
    a.t1 = a.two ? q - 1 : a.t0;
    a.l1 = a.two ? T : a.l0;
    return a;
}

__device__ inline AxisTiles axis_tangential(int i, int nt) {
    const int s = i + 1, q = s / T, r = s - q * T;
    AxisTiles a;
    a.t0 = q < nt ? q : q - 1;  // q == nt only for the last cell, r == 0
    a.l0 = q < nt ? r : r + T;
    a.two = r <= 1 && q >= 1 && q < nt;
    a.t1 = a.two ? q - 1 : a.t0;
    a.l1 = a.two ? r + T : a.l0;
    return a;
}

// Sum of one quantity pair (offsets OM, OU in the region; U only if has_u)
// over the <= 8 tile regions that hold one output, as exact 64-bit integers
// (the order of the adds does not matter).  Every load is issued before the
// first add: the pair along Z is read unconditionally (candidate 1 repeats
// candidate 0 when there is one tile), the extra X and Y tiles behind
// branches that are mostly uniform over a wave, and the values are pinned
// before the sums so that no load waits for another.
template <int SX, int SY, int OM, int OU>
__device__ inline void gather8(const unsigned long long* __restrict__ stage, unsigned R, const TileGeom& g,
                               const AxisTiles& X, const AxisTiles& Y, const AxisTiles& Z, bool has_u,
                               unsigned long long& am, unsigned long long& au) {
    auto off = [&](int tx, int lx, int ty, int ly) {
        return ((unsigned)(tx * g.nty + ty) * (unsigned)g.ntz) * R + (unsigned)(lx * SX + ly * SY);
    };
    const unsigned z0 = (unsigned)Z.t0 * R + (unsigned)Z.l0, z1 = (unsigned)Z.t1 * R + (unsigned)Z.l1;
    unsigned long long m[8] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
    unsigned long long u[8] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
    const unsigned b00 = off(X.t0, X.l0, Y.t0, Y.l0);
    m[0] = stage[b00 + z0 + OM];
    m[1] = stage[b00 + z1 + OM];
    if (has_u) {
        u[0] = stage[b00 + z0 + OU];
        u[1] = stage[b00 + z1 + OU];
    }
    if (Y.two) {
        const unsigned b = off(X.t0, X.l0, Y.t1, Y.l1);
        m[2] = stage[b + z0 + OM];
        m[3] = stage[b + z1 + OM];
        if (has_u) {
            u[2] = stage[b + z0 + OU];
            u[3] = stage[b + z1 + OU];
        }
    }
    if (X.two) {
        const unsigned b = off(X.t1, X.l1, Y.t0, Y.l0);
        m[4] = stage[b + z0 + OM];
        m[5] = stage[b + z1 + OM];
        if (has_u) {
            u[4] = stage[b + z0 + OU];
            u[5] = stage[b + z1 + OU];
        }
        if (Y.two) {
            const unsigned b2 = off(X.t1, X.l1, Y.t1, Y.l1);
            m[6] = stage[b2 + z0 + OM];
            m[7] = stage[b2 + z1 + OM];
            if (has_u) {
                u[6] = stage[b2 + z0 + OU];
                u[7] = stage[b2 + z1 + OU];
            }
        }
    }
    asm volatile("" ::"v"(m[0]), "v"(m[1]), "v"(m[2]), "v"(m[3]), "v"(m[4]), "v"(m[5]), "v"(m[6]), "v"(m[7]),
                 "v"(u[0]), "v"(u[1]), "v"(u[2]), "v"(u[3]), "v"(u[4]), "v"(u[5]), "v"(u[6]), "v"(u[7]));
    const unsigned long long zm = Z.two ? ~0ull : 0ull;
    am = m[0] + (m[1] & zm) + m[2] + (m[3] & zm) + m[4] + (m[5] & zm) + m[6] + (m[7] & zm);
    au = u[0] + (u[1] & zm) + u[2] + (u[3] & zm) + u[4] + (u[5] & zm) + u[6] + (u[7] & zm);
}

struct GatherArgs {
    TileGeom g;
    const unsigned long long* stage;
    int R;
    const float* bounds;
    const int* cmax;
    float* mass[3];
    float* mom[3];
    float* cmass;
    float* cwsum;
};

// One thread per point of the (nx+1)(ny+1)(nz+1) lattice: the three low
// faces and the cell of the point, each the sum of the tile regions that
// hold it.  32-bit index arithmetic: the entry point checks that the lattice
// and the staging buffer have fewer than 2^31 entries.
__global__ __launch_bounds__(256) void k_p2g_gather(GatherArgs a) {
    const TileGeom g = a.g;
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned ez = g.nz + 1, ey = g.ny + 1;
    if (t >= (unsigned)(g.nx + 1) * ey * ez) return;
    const FixScale fs = fix_scale(a.bounds, a.cmax);
    const unsigned row = t / ez;
    const int k = (int)(t - row * ez), i = (int)(row / ey), j = (int)(row - (unsigned)i * ey);
    const bool ci = i < g.nx, cj = j < g.ny, ck = k < g.nz;
    const unsigned R = (unsigned)a.R;
    // the tangential lists of an index past the last cell are never used
    const AxisTiles nX = axis_normal(i, g.ntx), nY = axis_normal(j, g.nty), nZ = axis_normal(k, g.ntz);
    const AxisTiles tX = axis_tangential(min(i, g.nx - 1), g.ntx), tY = axis_tangential(min(j, g.ny - 1), g.nty),
                    tZ = axis_tangential(min(k, g.nz - 1), g.ntz);
    const Dims d{g.nx, g.ny, g.nz};
    unsigned long long am, au;
    if (cj && ck) {
        gather8<E * E, E, O_XM, O_XU>(a.stage, R, g, nX, tY, tZ, true, am, au);
        const int f = face_index<0>(d, i, j, k);
        a.mass[0][f] = (float)((double)(long long)am * fs.im);
        a.mom[0][f] = (float)((double)(long long)au * fs.iu);
    }
    if (ci && ck) {
        gather8<F * E, E, O_YM, O_YU>(a.stage, R, g, tX, nY, tZ, true, am, au);
        const int f = face_index<1>(d, i, j, k);
        a.mass[1][f] = (float)((double)(long long)am * fs.im);
        a.mom[1][f] = (float)((double)(long long)au * fs.iu);
    }
    if (ci && cj) {
        gather8<E * F, F, O_ZM, O_ZU>(a.stage, R, g, tX, tY, nZ, true, am, au);
        const int f = face_index<2>(d, i, j, k);
        a.mass[2][f] = (float)((double)(long long)am * fs.im);
        a.mom[2][f] = (float)((double)(long long)au * fs.iu);
    }
    if (ci && cj && ck) {
        gather8<E * E, E, O_CM, O_CW>(a.stage, R, g, tX, tY, tZ, a.cwsum != nullptr, am, au);
        const int c = cell_index(d, i, j, k);
        a.cmass[c] = (float)((double)(long long)am * fs.im);
        if (a.cwsum) a.cwsum[c] = (float)((double)(long long)au * fs.iw);
    }
}

// ============================================================== G2P + advection
struct G2PArgs {
    float* px;
    float* pv;
    const int* ptype;
    const int* pstate;
    const uint8_t* pflag;
    int64_t n;
    Dims d;
    int flags;
    const float* vel[3];
    const float* dvel[3];
    const float* vel_liq[3];
    const int* solid;
    float dx, inv_dx, dt, alpha_l, alpha_a;
    float lo[3], hi[3];
};

// Trilinear interpolation of the face grid of axis A at grid coordinates
// g = x/dx, indices clamped per axis (kernels.py:52-79).  The two taps along
// k of each (i, j) row are adjacent in memory, so a sample is 4 row loads of
// 8 bytes: the pair (kc, kc+1) with kc = clamp(k0, 0, n2-2); where both taps
// clamp to one end of the row the pair supplies that end value twice.
template <int A>
struct Rows {
    int base[4];  // (ii * n1 + jj) * n2 + kc for (di, dj) = (0,0), (0,1), (1,0), (1,1)
    float w[8];   // wi * wj * wk, index di*4 + dj*2 + dk
    int mode;     // 0: taps (kc, kc+1); 1: both at 0; 2: both at n2-1
};

// Per coordinate the two cell splits a sample needs: the face grid of the
// same axis (offset 0) and the two others (offset 1/2).
struct Coord {
    int in, it;     // floor(g), floor(g - 1/2)
    float fn, ft;   // the fractions
};

__device__ inline Coord coord(float g) {
    Coord c;
    c.in = (int)floorf(g);
    c.fn = g - (float)c.in;
    const float h = g - 0.5f;
    c.it = (int)floorf(h);
    c.ft = h - (float)c.it;
    return c;
}

template <int A>
__device__ inline void make_rows(const Coord& cx, const Coord& cy, const Coord& cz, const Dims& d, Rows<A>& t) {
    const int n0 = d.nx + (A == 0), n1 = d.ny + (A == 1), n2 = d.nz + (A == 2);
    const int i0 = A == 0 ? cx.in : cx.it, j0 = A == 1 ? cy.in : cy.it, k0 = A == 2 ? cz.in : cz.it;
    const float fx = A == 0 ? cx.fn : cx.ft, fy = A == 1 ? cy.fn : cy.ft, fz = A == 2 ? cz.fn : cz.ft;
    const int kc = min(max(k0, 0), n2 - 2);
    t.mode = k0 < 0 ? 1 : (k0 > n2 - 2 ? 2 : 0);
#pragma unroll
    for (int di = 0; di < 2; ++di) {
        const float wi = di == 0 ? 1.0f - fx : fx;
        const int ii = min(max(i0 + di, 0), n0 - 1);
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const float wj = dj == 0 ? 1.0f - fy : fy;
            const int jj = min(max(j0 + dj, 0), n1 - 1);
            t.base[di * 2 + dj] = (ii * n1 + jj) * n2 + kc;
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                const float wk = dk == 0 ? 1.0f - fz : fz;
                t.w[di * 4 + dj * 2 + dk] = wi * wj * wk;
            }
        }
    }
}

__device__ inline float2 ld_row_pair(const float* __restrict__ f, int idx) {
    float2 v;
    __builtin_memcpy(&v, f + idx, sizeof(v));  // 4-byte aligned 8-byte load
    return v;
}

template <int A>
__device__ inline float sample_rows(const Rows<A>& t, const float* __restrict__ f) {
    float acc = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float2 v = ld_row_pair(f, t.base[r]);
        const float v0 = t.mode == 2 ? v.y : v.x;
        const float v1 = t.mode == 1 ? v.x : v.y;
        acc += t.w[2 * r] * v0;
        acc += t.w[2 * r + 1] * v1;
    }
    return acc;
}

struct Rows3 {
    Rows<0> t0;
    Rows<1> t1;
    Rows<2> t2;
};

__device__ inline void make_rows3(float x, float y, float z, float inv_dx, const Dims& d, Rows3& r) {
    const Coord cx = coord(x * inv_dx), cy = coord(y * inv_dx), cz = coord(z * inv_dx);
    make_rows(cx, cy, cz, d, r.t0);
    make_rows(cx, cy, cz, d, r.t1);
    make_rows(cx, cy, cz, d, r.t2);
}

// Two RK3 stages through the face fields f (uniform pointers: the caller
// branches on the particle's phase), from k1 at x (kernels.py:508-549).
__device__ inline void rk3_rest(const float* __restrict__ f0, const float* __restrict__ f1,
                                const float* __restrict__ f2, float& x, float& y, float& z, float k1x, float k1y,
                                float k1z, float dt, float inv_dx, const Dims& d) {
    const float h1 = 0.5f * dt, h2 = 0.75f * dt;
    Rows3 r;
    make_rows3(x + h1 * k1x, y + h1 * k1y, z + h1 * k1z, inv_dx, d, r);
    const float k2x = sample_rows(r.t0, f0), k2y = sample_rows(r.t1, f1), k2z = sample_rows(r.t2, f2);
    make_rows3(x + h2 * k2x, y + h2 * k2y, z + h2 * k2z, inv_dx, d, r);
    const float k3x = sample_rows(r.t0, f0), k3y = sample_rows(r.t1, f1), k3z = sample_rows(r.t2, f2);
    const float c1 = 2.0f / 9.0f, c2 = 3.0f / 9.0f, c3 = 4.0f / 9.0f;
    x = x + dt * (c1 * k1x + c2 * k2x + c3 * k3x);
    y = y + dt * (c1 * k1y + c2 * k2y + c3 * k3y);
    z = z + dt * (c1 * k1z + c2 * k2z + c3 * k3z);
}

// Wall reflection (kernels.py:567-595), solid push-out (kernels.py:630-679)
// and the stores, shared by the per-particle and the tiled kernels.
__device__ inline void finish_particle(const G2PArgs& a, int64_t p, float x, float y, float z, float& vx, float& vy,
                                       float& vz) {
    const Dims d = a.d;
    if (a.flags & PF_GA_REFLECT) {
        // kernels.py:567-595
        float xs[3] = {x, y, z}, vs[3] = {vx, vy, vz};
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            if (xs[ax] <= a.lo[ax]) {
                xs[ax] = a.lo[ax];
                vs[ax] = fmaxf(vs[ax], 0.f);
            } else if (xs[ax] >= a.hi[ax]) {
                xs[ax] = a.hi[ax];
                vs[ax] = fminf(vs[ax], 0.f);
            }
        }
        x = xs[0];
        y = xs[1];
        z = xs[2];
        vx = vs[0];
        vy = vs[1];
        vz = vs[2];
    }
    if ((a.flags & PF_GA_SOLIDS) && a.solid) {
        // kernels.py:630-679
        const int i = min(max((int)(x / a.dx), 0), d.nx - 1);
        const int j = min(max((int)(y / a.dx), 0), d.ny - 1);
        const int k = min(max((int)(z / a.dx), 0), d.nz - 1);
        if (a.solid[cell_index(d, i, j, k)] != 0) {
            float best = 1.0e9f;
            int bax = 0, bsg = 1;
            for (int dd = 0; dd < 6; ++dd) {
                const int ax = dd / 2, sg = (dd % 2 == 0) ? -1 : 1;
                for (int t = 1; t < 8; ++t) {
                    const int ii = i + (ax == 0 ? sg * t : 0);
                    const int jj = j + (ax == 1 ? sg * t : 0);
                    const int kk = k + (ax == 2 ? sg * t : 0);
                    if (ii < 0 || jj < 0 || kk < 0 || ii >= d.nx || jj >= d.ny || kk >= d.nz) break;
                    if (a.solid[cell_index(d, ii, jj, kk)] == 0) {
                        const float dist = (float)t;
                        if (dist < best) {
                            best = dist;
                            bax = ax;
                            bsg = sg;
                        }
                        break;
                    }
                }
            }
            if (best <= 1.0e8f) {
                const float off = (best + 0.5f) * a.dx * (float)bsg;
                float xs[3] = {x, y, z}, vs[3] = {vx, vy, vz};
                xs[bax] = xs[bax] + off;
                if ((float)bsg * vs[bax] < 0.f) vs[bax] = 0.f;
                x = xs[0];
                y = xs[1];
                z = xs[2];
                vx = vs[0];
                vy = vs[1];
                vz = vs[2];
            }
        }
    }
    a.px[3 * p] = x;
    a.px[3 * p + 1] = y;
    a.px[3 * p + 2] = z;
    a.pv[3 * p] = vx;
    a.pv[3 * p + 1] = vy;
    a.pv[3 * p + 2] = vz;
}

// At most 5 waves per SIMD: the kernel is bound by its 60 row-pair gathers
// per particle, and more resident waves evict each other's rows from L1/L2
// (measured at 128^3 x 27 ppc: 7 waves/SIMD is 10 % slower than 5).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 5))) void k_g2p_advect(G2PArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    bool liquid, active, dead;
    if (a.pflag) {
        const uint8_t fl = a.pflag[p];
        liquid = fl & 1;
        active = fl & 2;
        dead = fl & 4;
    } else {
        const int st = a.pstate[p];
        liquid = a.ptype[p] == 1;
        active = st == 0;
        dead = st == 2;
    }
    if (dead) return;
    float x = a.px[3 * p], y = a.px[3 * p + 1], z = a.px[3 * p + 2];
    float vx = a.pv[3 * p], vy = a.pv[3 * p + 1], vz = a.pv[3 * p + 2];
    const Dims d = a.d;
    const bool g2p = active && (a.flags & PF_GA_G2P);
    const bool adv = active && (a.flags & PF_GA_ADVECT);

    // samples at x share one set of rows: PIC I(vel), FLIP I(dvel), and the
    // first RK3 stage (liquid: I(vel_liq); air: I(vel) = the PIC sample)
    float k1x = 0.f, k1y = 0.f, k1z = 0.f;
    if (g2p || adv) {
        Rows3 r;
        make_rows3(x, y, z, a.inv_dx, d, r);
        float pic[3] = {0.f, 0.f, 0.f};
        if (g2p || !liquid) {
            pic[0] = sample_rows(r.t0, a.vel[0]);
            pic[1] = sample_rows(r.t1, a.vel[1]);
            pic[2] = sample_rows(r.t2, a.vel[2]);
        }
        if (adv) {
            if (liquid) {
                k1x = sample_rows(r.t0, a.vel_liq[0]);
                k1y = sample_rows(r.t1, a.vel_liq[1]);
                k1z = sample_rows(r.t2, a.vel_liq[2]);
            } else {
                k1x = pic[0];
                k1y = pic[1];
                k1z = pic[2];
            }
        }
        if (g2p) {
            // eq. (12): v = alpha (v + I(du)) + (1 - alpha) I(u)
            const float dux = sample_rows(r.t0, a.dvel[0]), duy = sample_rows(r.t1, a.dvel[1]),
                        duz = sample_rows(r.t2, a.dvel[2]);
            const float al = liquid ? a.alpha_l : a.alpha_a;
            vx = al * (vx + dux) + (1.0f - al) * pic[0];
            vy = al * (vy + duy) + (1.0f - al) * pic[1];
            vz = al * (vz + duz) + (1.0f - al) * pic[2];
        }
    }
    if (adv) {
        // Ralston RK3; liquid samples vel_liq, air vel
        if (liquid)
            rk3_rest(a.vel_liq[0], a.vel_liq[1], a.vel_liq[2], x, y, z, k1x, k1y, k1z, a.dt, a.inv_dx, d);
        else
            rk3_rest(a.vel[0], a.vel[1], a.vel[2], x, y, z, k1x, k1y, k1z, a.dt, a.inv_dx, d);
        x = fminf(fmaxf(x, a.lo[0]), a.hi[0]);
        y = fminf(fmaxf(y, a.lo[1]), a.hi[1]);
        z = fminf(fmaxf(z, a.lo[2]), a.hi[2]);
    }
    finish_particle(a, p, x, y, z, vx, vy, vz);
}

// Per-particle terms of the CFL speed and the P2G fixed-point bounds: |v|,
// |m|, |m| max_a |v_a| (particle_max_speed, kernels.py:552-564).  One
// definition for both kernels that produce them, so the maxima agree bit for
// bit.
struct SpeedBounds {
    float v = 0.f, m = 0.f, mu = 0.f;
    __device__ inline void add(float vx, float vy, float vz, float mass) {
        v = fmaxf(v, sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(vx, vx), __fmul_rn(vy, vy)), __fmul_rn(vz, vz))));
        const float ma = fabsf(mass);
        m = fmaxf(m, ma);
        mu = fmaxf(mu, __fmul_rn(ma, fmaxf(fabsf(vx), fmaxf(fabsf(vy), fabsf(vz)))));
    }
    // block maxima into out[0..2] (non-negative floats, zeroed beforehand)
    __device__ inline void flush(float* out, float* sh) {
        const float a = block_max(v, sh);
        if (threadIdx.x == 0) atomic_max_nonneg(out, a);
        const float b = block_max(m, sh);
        if (threadIdx.x == 0) atomic_max_nonneg(out + 1, b);
        const float c = block_max(mu, sh);
        if (threadIdx.x == 0) atomic_max_nonneg(out + 2, c);
    }
};

// ============================================================== tiled G2P + advection
// One workgroup per tile of T^3 cells.  The face grids the tile's particles
// sample (vel, dvel, vel_liq) are staged in LDS over the tile plus a 2-cell
// halo on every side (GB = T + 4 entries per axis), indices clamped into each
// grid while staging: an LDS tap then holds exactly what the reference's
// clamped tap reads, so every sample is the reference's sum in the
// reference's order with no per-tap clamps.  The staged box covers every
// sample within one cell of the tile (the first RK3 stage always, the later
// stages at CFL <= 1); a sample outside it reads global memory.
constexpr int GB = T + 4;
constexpr int GB2 = GB * GB;
constexpr int GB3 = GB * GB * GB;
constexpr int kG2PThreads = 512;

struct BoxRows {
    int base;    // LDS offset of tap (0,0,0)
    float w[8];  // wi * wj * wk, index di*4 + dj*2 + dk
};

template <int A>
__device__ inline bool box_rows(const Coord& cx, const Coord& cy, const Coord& cz, int ox, int oy, int oz,
                                BoxRows& r) {
    const int li = (A == 0 ? cx.in : cx.it) - ox;
    const int lj = (A == 1 ? cy.in : cy.it) - oy;
    const int lk = (A == 2 ? cz.in : cz.it) - oz;
    const float fx = A == 0 ? cx.fn : cx.ft, fy = A == 1 ? cy.fn : cy.ft, fz = A == 2 ? cz.fn : cz.ft;
    r.base = (li * GB + lj) * GB + lk;
#pragma unroll
    for (int di = 0; di < 2; ++di) {
        const float wi = di == 0 ? 1.0f - fx : fx;
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const float wj = dj == 0 ? 1.0f - fy : fy;
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) r.w[di * 4 + dj * 2 + dk] = wi * wj * (dk == 0 ? 1.0f - fz : fz);
        }
    }
    return (unsigned)li <= (unsigned)(GB - 2) && (unsigned)lj <= (unsigned)(GB - 2) &&
           (unsigned)lk <= (unsigned)(GB - 2);
}

__device__ inline float sample_box(const BoxRows& r, const float* L) {
    const float* q = L + r.base;
    float acc = 0.f;
    acc += r.w[0] * q[0];
    acc += r.w[1] * q[1];
    acc += r.w[2] * q[GB];
    acc += r.w[3] * q[GB + 1];
    acc += r.w[4] * q[GB2];
    acc += r.w[5] * q[GB2 + 1];
    acc += r.w[6] * q[GB2 + GB];
    acc += r.w[7] * q[GB2 + GB + 1];
    return acc;
}

struct BoxRows3 {
    BoxRows r0, r1, r2;
};

__device__ inline bool box_rows3(float x, float y, float z, float inv_dx, int ox, int oy, int oz, BoxRows3& b) {
    const Coord cx = coord(x * inv_dx), cy = coord(y * inv_dx), cz = coord(z * inv_dx);
    const bool i0 = box_rows<0>(cx, cy, cz, ox, oy, oz, b.r0);
    const bool i1 = box_rows<1>(cx, cy, cz, ox, oy, oz, b.r1);
    const bool i2 = box_rows<2>(cx, cy, cz, ox, oy, oz, b.r2);
    return i0 && i1 && i2;
}

// Velocity sample at (x, y, z) from the staged field `fb` (LDS, 3 axes of
// GB3 entries each) or, outside the box, from the global fields `fg`.
__device__ inline void sample3(const float* fb, const float* const* fg, float x, float y, float z,
                               const G2PArgs& a, int ox, int oy, int oz, float& u, float& v, float& w) {
    BoxRows3 b;
    if (box_rows3(x, y, z, a.inv_dx, ox, oy, oz, b)) {
        u = sample_box(b.r0, fb);
        v = sample_box(b.r1, fb + GB3);
        w = sample_box(b.r2, fb + 2 * GB3);
    } else {
        Rows3 r;
        make_rows3(x, y, z, a.inv_dx, a.d, r);
        u = sample_rows(r.t0, fg[0]);
        v = sample_rows(r.t1, fg[1]);
        w = sample_rows(r.t2, fg[2]);
    }
}

// One face array's part of the box: the (clamped) grid index of box entry
// e, e in [0, GB3).
template <int AX>
__device__ inline int64_t box_src_index(int e, const Dims& d, int ox, int oy, int oz) {
    const int n0 = d.nx + (AX == 0), n1 = d.ny + (AX == 1), n2 = d.nz + (AX == 2);
    const int li = e / GB2, lj = (e / GB) % GB, lk = e % GB;
    const int gi = min(max(ox + li, 0), n0 - 1);
    const int gj = min(max(oy + lj, 0), n1 - 1);
    const int gk = min(max(oz + lk, 0), n2 - 1);
    return ((int64_t)gi * n1 + gj) * n2 + gk;
}

// Stage the fields of the box (vel; dvel with g2p; vel_liq with adv): every
// load of a thread is issued before the first LDS store, so the staging
// costs one round trip to memory instead of one per element.
constexpr int kStageIters = (GB3 + kG2PThreads - 1) / kG2PThreads;
__device__ inline void stage_fields(float* box, const G2PArgs& a, bool g2p, bool adv, int ox, int oy, int oz) {
    float v[9][kStageIters];
#pragma unroll
    for (int q = 0; q < kStageIters; ++q) {
        // the last pass's idle threads read entry GB3-1 again (not stored)
        const int e = min((int)threadIdx.x + q * kG2PThreads, GB3 - 1);
        const int64_t i0 = box_src_index<0>(e, a.d, ox, oy, oz), i1 = box_src_index<1>(e, a.d, ox, oy, oz),
                      i2 = box_src_index<2>(e, a.d, ox, oy, oz);
        v[0][q] = a.vel[0][i0];
        v[1][q] = a.vel[1][i1];
        v[2][q] = a.vel[2][i2];
        if (g2p) {
            v[3][q] = a.dvel[0][i0];
            v[4][q] = a.dvel[1][i1];
            v[5][q] = a.dvel[2][i2];
        }
        if (adv) {
            v[6][q] = a.vel_liq[0][i0];
            v[7][q] = a.vel_liq[1][i1];
            v[8][q] = a.vel_liq[2][i2];
        }
    }
#pragma unroll
    for (int q = 0; q < kStageIters; ++q) {
        const int e = (int)threadIdx.x + q * kG2PThreads;
        if (e < GB3) {
#pragma unroll
            for (int f = 0; f < 3; ++f) box[f * GB3 + e] = v[f][q];
            if (g2p)
#pragma unroll
                for (int f = 3; f < 6; ++f) box[f * GB3 + e] = v[f][q];
            if (adv)
#pragma unroll
                for (int f = 6; f < 9; ++f) box[f * GB3 + e] = v[f][q];
        }
    }
}

// bounds (optional, with pmass): the CFL speed and P2G bounds of the
// particles after the step, as pf_max_speed would compute them next step.
__global__ __launch_bounds__(kG2PThreads) void k_g2p_tiled(G2PArgs a, const int* __restrict__ perm,
                                                           const int* __restrict__ start, int ntx, int nty,
                                                           int ntz, const float* __restrict__ pmass,
                                                           float* __restrict__ bounds) {
    // fields: [0, 3 GB3) vel, [3 GB3, 6 GB3) dvel, [6 GB3, 9 GB3) vel_liq
    __shared__ float box[9 * GB3];
    const int nt = ntx * nty * ntz;
    const int t = xcd_block(blockIdx.x, nt);
    const int tz = t % ntz, ty = (t / ntz) % nty, tx = t / (ntz * nty);
    const int ox = tx * T - 2, oy = ty * T - 2, oz = tz * T - 2;
    const int s0 = start[(int64_t)t * T3], s1 = start[(int64_t)(t + 1) * T3];
    if (s0 == s1) return;
    const bool g2p = a.flags & PF_GA_G2P, adv = a.flags & PF_GA_ADVECT;
    SpeedBounds sb;
    stage_fields(box, a, g2p, adv, ox, oy, oz);
    __syncthreads();

    // the next particle's record is loaded before the current one is
    // processed, so its latency hides behind that work
    int s = s0 + (int)threadIdx.x;
    int64_t pn = 0;
    bool liq_n = false;
    float xn = 0.f, yn = 0.f, zn = 0.f, vxn = 0.f, vyn = 0.f, vzn = 0.f;
    if (s < s1) {
        pn = perm[s];
        liq_n = a.pflag ? (a.pflag[pn] & 1) : (a.ptype[pn] == 1);
        xn = a.px[3 * pn], yn = a.px[3 * pn + 1], zn = a.px[3 * pn + 2];
        vxn = a.pv[3 * pn], vyn = a.pv[3 * pn + 1], vzn = a.pv[3 * pn + 2];
    }
    for (; s < s1; s += kG2PThreads) {
        const int64_t p = pn;
        const bool liquid = liq_n;
        float x = xn, y = yn, z = zn;
        float vx = vxn, vy = vyn, vz = vzn;
        if (s + kG2PThreads < s1) {
            pn = perm[s + kG2PThreads];
            liq_n = a.pflag ? (a.pflag[pn] & 1) : (a.ptype[pn] == 1);
            xn = a.px[3 * pn], yn = a.px[3 * pn + 1], zn = a.px[3 * pn + 2];
            vxn = a.pv[3 * pn], vyn = a.pv[3 * pn + 1], vzn = a.pv[3 * pn + 2];
        }
        const float* fk = liquid ? box + 6 * GB3 : box;                 // RK3 field in LDS
        const float* const* gk = liquid ? a.vel_liq : a.vel;            // and in HBM
        float k1x = 0.f, k1y = 0.f, k1z = 0.f;
        {
            BoxRows3 b;
            if (box_rows3(x, y, z, a.inv_dx, ox, oy, oz, b)) {
                float pic[3] = {0.f, 0.f, 0.f};
                if (g2p || !liquid) {
                    pic[0] = sample_box(b.r0, box);
                    pic[1] = sample_box(b.r1, box + GB3);
                    pic[2] = sample_box(b.r2, box + 2 * GB3);
                }
                if (adv) {
                    if (liquid) {
                        k1x = sample_box(b.r0, box + 6 * GB3);
                        k1y = sample_box(b.r1, box + 7 * GB3);
                        k1z = sample_box(b.r2, box + 8 * GB3);
                    } else {
                        k1x = pic[0];
                        k1y = pic[1];
                        k1z = pic[2];
                    }
                }
                if (g2p) {
                    const float dux = sample_box(b.r0, box + 3 * GB3), duy = sample_box(b.r1, box + 4 * GB3),
                                duz = sample_box(b.r2, box + 5 * GB3);
                    const float al = liquid ? a.alpha_l : a.alpha_a;
                    vx = al * (vx + dux) + (1.0f - al) * pic[0];
                    vy = al * (vy + duy) + (1.0f - al) * pic[1];
                    vz = al * (vz + duz) + (1.0f - al) * pic[2];
                }
            } else {
                // a particle outside its tile's box (a position the binning
                // clamped into the grid): the per-particle path's samples
                Rows3 r;
                make_rows3(x, y, z, a.inv_dx, a.d, r);
                float pic[3] = {0.f, 0.f, 0.f};
                if (g2p || !liquid) {
                    pic[0] = sample_rows(r.t0, a.vel[0]);
                    pic[1] = sample_rows(r.t1, a.vel[1]);
                    pic[2] = sample_rows(r.t2, a.vel[2]);
                }
                if (adv) {
                    if (liquid) {
                        k1x = sample_rows(r.t0, a.vel_liq[0]);
                        k1y = sample_rows(r.t1, a.vel_liq[1]);
                        k1z = sample_rows(r.t2, a.vel_liq[2]);
                    } else {
                        k1x = pic[0];
                        k1y = pic[1];
                        k1z = pic[2];
                    }
                }
                if (g2p) {
                    const float dux = sample_rows(r.t0, a.dvel[0]), duy = sample_rows(r.t1, a.dvel[1]),
                                duz = sample_rows(r.t2, a.dvel[2]);
                    const float al = liquid ? a.alpha_l : a.alpha_a;
                    vx = al * (vx + dux) + (1.0f - al) * pic[0];
                    vy = al * (vy + duy) + (1.0f - al) * pic[1];
                    vz = al * (vz + duz) + (1.0f - al) * pic[2];
                }
            }
        }
        if (adv) {
            // Ralston RK3 (kernels.py:508-549)
            const float dt = a.dt, h1 = 0.5f * dt, h2 = 0.75f * dt;
            float k2x, k2y, k2z, k3x, k3y, k3z;
            sample3(fk, gk, x + h1 * k1x, y + h1 * k1y, z + h1 * k1z, a, ox, oy, oz, k2x, k2y, k2z);
            sample3(fk, gk, x + h2 * k2x, y + h2 * k2y, z + h2 * k2z, a, ox, oy, oz, k3x, k3y, k3z);
            const float c1 = 2.0f / 9.0f, c2 = 3.0f / 9.0f, c3 = 4.0f / 9.0f;
            x = x + dt * (c1 * k1x + c2 * k2x + c3 * k3x);
            y = y + dt * (c1 * k1y + c2 * k2y + c3 * k3y);
            z = z + dt * (c1 * k1z + c2 * k2z + c3 * k3z);
            x = fminf(fmaxf(x, a.lo[0]), a.hi[0]);
            y = fminf(fmaxf(y, a.lo[1]), a.hi[1]);
            z = fminf(fmaxf(z, a.lo[2]), a.hi[2]);
        }
        finish_particle(a, p, x, y, z, vx, vy, vz);
        if (bounds) sb.add(vx, vy, vz, pmass[p]);
    }
    if (bounds) {
        __shared__ float sh[kG2PThreads / 64];
        sb.flush(bounds, sh);
    }
}

// Particles outside the bins (pstate != 0 and not dead, i.e. droplets) after
// the tiled pass: wall reflection and solids only, as in the per-particle
// kernel.  Exits at once when the bins held every particle.
__global__ __launch_bounds__(256) void k_g2p_unbinned(G2PArgs a, const int* __restrict__ n_binned) {
    if (*n_binned == a.n) return;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.n;
         p += (int64_t)gridDim.x * blockDim.x) {
        bool active, dead;
        if (a.pflag) {
            const uint8_t fl = a.pflag[p];
            active = fl & 2;
            dead = fl & 4;
        } else {
            const int st = a.pstate[p];
            active = st == 0;
            dead = st == 2;
        }
        if (active || dead) continue;
        float vx = a.pv[3 * p], vy = a.pv[3 * p + 1], vz = a.pv[3 * p + 2];
        finish_particle(a, p, a.px[3 * p], a.px[3 * p + 1], a.px[3 * p + 2], vx, vy, vz);
    }
}

// ============================================================== CFL speed + splat bounds
// out[0] = max |v| over FLIP particles (particle_max_speed, kernels.py:552-564);
// with pmass: out[1] = max m, out[2] = max m * max_a |v_a| (bounds that size
// the fixed-point P2G accumulators).  Maxima of non-negative floats through
// integer atomics: order-independent.
__global__ __launch_bounds__(256) void k_max_speed(const float* __restrict__ pv,
                                                   const float* __restrict__ pmass,
                                                   const int* __restrict__ pstate, int64_t n,
                                                   float* out) {
    __shared__ float sh[4];
    SpeedBounds sb;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        if (pstate[p] != 0) continue;
        sb.add(pv[3 * p], pv[3 * p + 1], pv[3 * p + 2], pmass ? pmass[p] : 0.0f);
    }
    if (pmass) {
        sb.flush(out, sh);
    } else {
        const float m = block_max(sb.v, sh);
        if (threadIdx.x == 0) atomic_max_nonneg(out, m);
    }
}

}  // namespace pf

using namespace pf;

// ============================================================== entry points
PF_API int64_t pf_bin_nkeys(int nx, int ny, int nz) {
    return n_tiles(tile_geom(nx, ny, nz)) * T3;
}

PF_API size_t pf_bin_scan_tmp_bytes(int nx, int ny, int nz) {
    const int64_t nb = pf_bin_nkeys(nx, ny, nz) + 1;  // cells + the inactive bucket
    // count array + block sums + offsets
    return (size_t)(nb + 2 * (cdiv(nb + 1, SCAN_TILE) + 2)) * sizeof(int);
}

// Argument checks shared by pf_bin_sort and pf_bin_sort_noalloc (everything
// but the scan buffer).
static int bin_sort_check(const char* fn, const float* px, const int32_t* pstate, int64_t n, int nx,
                          int ny, int nz, float dx, const uint32_t* order, const uint32_t* keys,
                          const uint32_t* rank, const uint32_t* perm, const int32_t* cell_start) {
    PF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && n >= 0, "%s: bad sizes", fn);
    PF_REQUIRE(n < (int64_t)1 << 31, "%s: more than 2^31 particles", fn);
    PF_REQUIRE((int64_t)nx * ny * nz < (int64_t)1 << 31, "%s: grid too large", fn);
    PF_REQUIRE(!(n > 0 && (!px || !pstate || !keys || !rank || !perm)), "%s: null particle buffer", fn);
    PF_REQUIRE(cell_start, "%s: null cell_start", fn);
    PF_REQUIRE(dx > 0.f, "%s: dx must be positive", fn);
    PF_REQUIRE(order != perm || n == 0, "%s: order and perm must be distinct buffers", fn);
    return PF_OK;
}

// The sort itself, on a scan buffer of pf_bin_scan_tmp_bytes() bytes.
static int bin_sort_run(const float* px, const int32_t* pstate, int64_t n, int nx, int ny, int nz,
                        float dx, const uint32_t* order, uint32_t* keys, uint32_t* rank,
                        uint32_t* perm, int32_t* cell_start, void* scan_tmp, int32_t* n_active,
                        int32_t stable, hipStream_t s) {
    const TileGeom g = tile_geom(nx, ny, nz);
    const int64_t nkeys = n_tiles(g) * T3;
    const int64_t nbuckets = nkeys + 1;
    int* count = static_cast<int*>(scan_tmp);
    int* tmp = count + nbuckets;
    int rc = fill_u32(count, nbuckets, 0u, s);
    if (rc) return rc;
    if (n > 0) {
        k_bin<<<cdiv(n, 256), 256, 0, s>>>(px, pstate, n, order, g, dx, (uint32_t)nkeys, keys, rank,
                                           count);
        PF_CHECK_LAUNCH();
    }
    rc = exclusive_scan(count, nbuckets, cell_start, tmp, s);
    if (rc) return rc;
    if (n_active) {
        k_copy_count<<<1, 64, 0, s>>>(cell_start, nkeys, n_active);
        PF_CHECK_LAUNCH();
    }
    if (n > 0) {
        k_scatter<<<cdiv(n, 256), 256, 0, s>>>(keys, rank, n, order, (uint32_t)nkeys, cell_start, perm);
        PF_CHECK_LAUNCH();
        if (stable) {
            k_key_sort<<<(unsigned)(nkeys / SLAB_KEYS), 64, 0, s>>>(cell_start, perm);
            PF_CHECK_LAUNCH();
        }
    }
    return PF_OK;
}

PF_API int pf_bin_sort(const float* px, const int32_t* pstate, int64_t n, int nx, int ny, int nz,
                       float dx, const uint32_t* order, uint32_t* keys, uint32_t* rank,
                       uint32_t* perm, int32_t* cell_start, void* scan_tmp, int32_t* n_active,
                       int32_t stable, pf_stream_t stream) {
    int rc = bin_sort_check("pf_bin_sort", px, pstate, n, nx, ny, nz, dx, order, keys, rank, perm,
                            cell_start);
    if (rc) return rc;
    PF_REQUIRE(scan_tmp, "pf_bin_sort: null scan_tmp");
    return bin_sort_run(px, pstate, n, nx, ny, nz, dx, order, keys, rank, perm, cell_start, scan_tmp,
                        n_active, stable, S(stream));
}

// ------------------------------------------------ pf_bin_sort without allocation
// The scan buffer of pf_bin_sort_noalloc lives in this file-scope device
// array: one copy per GPU, held until the process ends, sized for
// PF_BIN_NOALLOC_MAX_KEYS binning keys.  In host code `scratch` names a
// host-side stand-in, not the device copy: kernels only ever receive the
// address hipGetSymbolAddress returns for the current device.
// (the layout of pf_bin_scan_tmp_bytes: counts of nkeys + 1 buckets, then the
// block sums and offsets of their scan)
constexpr int64_t kNoallocBuckets = (int64_t)PF_BIN_NOALLOC_MAX_KEYS + 1;
constexpr int64_t kNoallocScanInts =
    kNoallocBuckets + 2 * ((kNoallocBuckets + 1 + SCAN_TILE - 1) / SCAN_TILE + 2);
static __device__ __align__(256) int scratch[kNoallocScanInts];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

// The device address of the current device's `scratch`, looked up once per
// device (with that device current) and kept.
static int noalloc_scratch(void** out) {
    int dev = 0;
    PF_CHECK_HIP(hipGetDevice(&dev));
    PF_REQUIRE(dev >= 0 && dev < kMaxDevices, "pf_bin_sort_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        PF_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        PF_REQUIRE(ptr, "pf_bin_sort_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return PF_OK;
}

PF_API int pf_bin_sort_noalloc(const float* px, const int32_t* pstate, int64_t n, int nx, int ny,
                               int nz, float dx, const uint32_t* order, uint32_t* keys,
                               uint32_t* rank, uint32_t* perm, int32_t* cell_start,
                               int32_t* n_active, int32_t stable, pf_stream_t stream) {
    int rc = bin_sort_check("pf_bin_sort_noalloc", px, pstate, n, nx, ny, nz, dx, order, keys, rank,
                            perm, cell_start);
    if (rc) return rc;
    const int64_t nkeys = pf_bin_nkeys(nx, ny, nz);
    if (nkeys > PF_BIN_NOALLOC_MAX_KEYS)
        return set_error(PF_ERR_SIZE,
                         "pf_bin_sort_noalloc: %lld binning keys for a %dx%dx%d grid exceed the "
                         "built-in scratch of %d keys",
                         (long long)nkeys, nx, ny, nz, PF_BIN_NOALLOC_MAX_KEYS);
    void* tmp = nullptr;
    rc = noalloc_scratch(&tmp);
    if (rc) return rc;
    return bin_sort_run(px, pstate, n, nx, ny, nz, dx, order, keys, rank, perm, cell_start, tmp,
                        n_active, stable, S(stream));
}

PF_API size_t pf_p2g_stage_bytes(int nx, int ny, int nz, int with_wsum) {
    const TileGeom g = tile_geom(nx, ny, nz);
    // 64-bit fixed-point tile regions, then one int (max count per key)
    return (size_t)n_tiles(g) * (with_wsum ? REG_W : REG) * sizeof(unsigned long long) + 256;
}

PF_API int pf_p2g(const float* px, const float* pv, const float* pmass, const uint32_t* perm,
                  const int32_t* cell_start, int nx, int ny, int nz, float dx, float rp,
                  const float* bounds, float* mass_x, float* mass_y, float* mass_z, float* mom_x,
                  float* mom_y, float* mom_z, float* cell_mass, float* cell_wsum, void* stage,
                  size_t stage_bytes, pf_stream_t stream) {
    PF_REQUIRE(nx > 0 && ny > 0 && nz > 0, "pf_p2g: bad grid");
    PF_REQUIRE((int64_t)(nx + 1) * (ny + 1) * (nz + 1) < (int64_t)1 << 31, "pf_p2g: grid too large");
    PF_REQUIRE(dx > 0.f && rp > 0.f, "pf_p2g: dx and rp must be positive");
    // the tiled kernel visits the 2 candidate planes per axis that a support
    // radius of one cell can reach (kernels.py:137-146 with R = rp/dx = 1)
    PF_REQUIRE(rp == dx, "pf_p2g: the tiled splat supports r_p == dx only (got rp=%g dx=%g)",
               (double)rp, (double)dx);
    PF_REQUIRE(mass_x && mass_y && mass_z && mom_x && mom_y && mom_z && cell_mass,
               "pf_p2g: null grid output");
    PF_REQUIRE(perm && cell_start && stage && bounds, "pf_p2g: null perm / cell_start / stage / bounds");
    PF_REQUIRE(stage_bytes >= pf_p2g_stage_bytes(nx, ny, nz, cell_wsum != nullptr),
               "pf_p2g: stage buffer too small (%zu < %zu bytes)", stage_bytes,
               pf_p2g_stage_bytes(nx, ny, nz, cell_wsum != nullptr));
    if (n_tiles(tile_geom(nx, ny, nz)) * REG_W >= (int64_t)1 << 31)
        return set_error(PF_ERR_SIZE, "pf_p2g: grid too large (staging offsets must fit in 31 bits)");
    hipStream_t s = S(stream);
    P2GArgs a;
    a.px = px;
    a.pv = pv;
    a.pmass = pmass;
    a.perm = perm;
    a.start = cell_start;
    a.g = tile_geom(nx, ny, nz);
    a.dx = dx;
    a.inv_r2 = 1.0f / (rp * rp);
    a.stage = static_cast<unsigned long long*>(stage);
    a.bounds = bounds;
    const int64_t nt = n_tiles(a.g);
    const int R = cell_wsum ? REG_W : REG;
    int* cmax = reinterpret_cast<int*>(a.stage + nt * R);
    a.cmax = cmax;
    int rc = fill_u32(cmax, 1, 0u, s);
    if (rc) return rc;
    const int64_t nkeys = nt * T3;
    const int64_t nbk = cdiv(nkeys, 256 * 8);
    k_max_count<<<(unsigned)(nbk < 2048 ? nbk : 2048), 256, 0, s>>>(cell_start, nkeys, cmax);
    PF_CHECK_LAUNCH();
    if (cell_wsum)
        k_p2g<true><<<(unsigned)nt, kP2GThreads, 0, s>>>(a);
    else
        k_p2g<false><<<(unsigned)nt, kP2GThreads, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    GatherArgs ga;
    ga.g = a.g;
    ga.stage = a.stage;
    ga.R = R;
    ga.bounds = bounds;
    ga.cmax = cmax;
    ga.mass[0] = mass_x;
    ga.mass[1] = mass_y;
    ga.mass[2] = mass_z;
    ga.mom[0] = mom_x;
    ga.mom[1] = mom_y;
    ga.mom[2] = mom_z;
    ga.cmass = cell_mass;
    ga.cwsum = cell_wsum;
    const int64_t n = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    k_p2g_gather<<<cdiv(n, 256), 256, 0, s>>>(ga);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_g2p_advect(const pf_g2p_args* args, pf_stream_t stream) {
    PF_REQUIRE(args, "pf_g2p_advect: null args");
    const pf_g2p_args& in = *args;
    PF_REQUIRE(in.nx > 0 && in.ny > 0 && in.nz > 0 && in.n >= 0, "pf_g2p_advect: bad sizes");
    PF_REQUIRE(in.nz >= 2, "pf_g2p_advect: nz >= 2 is required (row-pair samples)");
    PF_REQUIRE(in.n < (int64_t)1 << 31, "pf_g2p_advect: more than 2^31 particles");
    PF_REQUIRE(in.dx > 0.f, "pf_g2p_advect: dx must be positive");
    if (in.n == 0) return PF_OK;
    PF_REQUIRE(in.px && in.pv, "pf_g2p_advect: null particles");
    PF_REQUIRE(in.pflag || (in.ptype && in.pstate), "pf_g2p_advect: need pflag or ptype+pstate");
    if (in.flags & PF_GA_G2P)
        for (int q = 0; q < 3; ++q)
            PF_REQUIRE(in.vel[q] && in.dvel[q], "pf_g2p_advect: G2P needs vel and dvel");
    if (in.flags & PF_GA_ADVECT)
        for (int q = 0; q < 3; ++q)
            PF_REQUIRE(in.vel[q] && in.vel_liq[q], "pf_g2p_advect: advection needs vel, vel_liq");
    PF_REQUIRE(!(in.flags & PF_GA_SOLIDS) || in.solid, "pf_g2p_advect: solids flag without mask");
    G2PArgs a;
    a.px = in.px;
    a.pv = in.pv;
    a.ptype = in.ptype;
    a.pstate = in.pstate;
    a.pflag = in.pflag;
    a.n = in.n;
    a.d = Dims{in.nx, in.ny, in.nz};
    a.flags = in.flags;
    for (int q = 0; q < 3; ++q) {
        a.vel[q] = in.vel[q];
        a.dvel[q] = in.dvel[q];
        a.vel_liq[q] = in.vel_liq[q];
        a.lo[q] = in.lo[q];
        a.hi[q] = in.hi[q];
    }
    a.solid = in.solid;
    a.dx = in.dx;
    // grid coordinates are x * (1/dx): the same bits as x / dx when dx is a
    // power of two (every grid here), within an ulp otherwise
    a.inv_dx = 1.0f / in.dx;
    a.dt = in.dt;
    a.alpha_l = in.alpha_liquid;
    a.alpha_a = in.alpha_air;
    k_g2p_advect<<<cdiv(in.n, 256), 256, 0, S(stream)>>>(a);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_g2p_advect_binned(const pf_g2p_args* args, const int32_t* perm, const int32_t* cell_start,
                                const int32_t* n_active, const float* pmass, float* bounds_out,
                                pf_stream_t stream) {
    PF_REQUIRE(args, "pf_g2p_advect_binned: null args");
    PF_REQUIRE(!bounds_out || pmass, "pf_g2p_advect_binned: bounds need pmass");
    const pf_g2p_args& in = *args;
    PF_REQUIRE(perm && cell_start && n_active, "pf_g2p_advect_binned: null bins");
    PF_REQUIRE(in.nx > 0 && in.ny > 0 && in.nz > 0 && in.n >= 0, "pf_g2p_advect_binned: bad sizes");
    PF_REQUIRE(in.nz >= 2, "pf_g2p_advect_binned: nz >= 2 is required (row-pair samples)");
    PF_REQUIRE(in.n < (int64_t)1 << 31, "pf_g2p_advect_binned: more than 2^31 particles");
    PF_REQUIRE((int64_t)in.nx * in.ny * in.nz < (int64_t)1 << 31, "pf_g2p_advect_binned: grid too large");
    PF_REQUIRE(in.dx > 0.f, "pf_g2p_advect_binned: dx must be positive");
    if (in.n == 0) return PF_OK;
    PF_REQUIRE(in.px && in.pv, "pf_g2p_advect_binned: null particles");
    PF_REQUIRE(in.pflag || (in.ptype && in.pstate), "pf_g2p_advect_binned: need pflag or ptype+pstate");
    PF_REQUIRE(in.flags & (PF_GA_G2P | PF_GA_ADVECT), "pf_g2p_advect_binned: nothing to do without G2P/ADVECT");
    for (int q = 0; q < 3; ++q) {
        PF_REQUIRE(in.vel[q], "pf_g2p_advect_binned: null vel");
        PF_REQUIRE(!(in.flags & PF_GA_G2P) || in.dvel[q], "pf_g2p_advect_binned: G2P needs dvel");
        PF_REQUIRE(!(in.flags & PF_GA_ADVECT) || in.vel_liq[q], "pf_g2p_advect_binned: advection needs vel_liq");
    }
    PF_REQUIRE(!(in.flags & PF_GA_SOLIDS) || in.solid, "pf_g2p_advect_binned: solids flag without mask");
    G2PArgs a;
    a.px = in.px;
    a.pv = in.pv;
    a.ptype = in.ptype;
    a.pstate = in.pstate;
    a.pflag = in.pflag;
    a.n = in.n;
    a.d = Dims{in.nx, in.ny, in.nz};
    a.flags = in.flags;
    for (int q = 0; q < 3; ++q) {
        a.vel[q] = in.vel[q];
        a.dvel[q] = in.dvel[q];
        a.vel_liq[q] = in.vel_liq[q];
        a.lo[q] = in.lo[q];
        a.hi[q] = in.hi[q];
    }
    a.solid = in.solid;
    a.dx = in.dx;
    a.inv_dx = 1.0f / in.dx;
    a.dt = in.dt;
    a.alpha_l = in.alpha_liquid;
    a.alpha_a = in.alpha_air;
    const TileGeom g = tile_geom(in.nx, in.ny, in.nz);
    const int64_t nt = n_tiles(g);
    if (bounds_out) {
        const int rc = fill_u32(bounds_out, 3, 0u, S(stream));
        if (rc) return rc;
    }
    k_g2p_tiled<<<(unsigned)nt, kG2PThreads, 0, S(stream)>>>(a, perm, cell_start, g.ntx, g.nty, g.ntz, pmass,
                                                             bounds_out);
    PF_CHECK_LAUNCH();
    // particles outside the bins (droplets): reflection and solids only
    if (in.flags & (PF_GA_REFLECT | PF_GA_SOLIDS)) {
        const int64_t nb = cdiv(in.n, 256);
        k_g2p_unbinned<<<(unsigned)(nb < 2048 ? nb : 2048), 256, 0, S(stream)>>>(a, n_active);
        PF_CHECK_LAUNCH();
    }
    return PF_OK;
}

PF_API int pf_max_speed(const float* pv, const float* pmass, const int32_t* pstate, int64_t n,
                        float* out, pf_stream_t stream) {
    PF_REQUIRE(out, "pf_max_speed: null output");
    PF_REQUIRE(n == 0 || (pv && pstate), "pf_max_speed: null particles");
    hipStream_t s = S(stream);
    int rc = fill_u32(out, pmass ? 3 : 1, 0u, s);
    if (rc) return rc;
    if (n == 0) return PF_OK;
    const int64_t nb = cdiv(n, 256 * 8);
    k_max_speed<<<(unsigned)(nb < 4096 ? nb : 4096), 256, 0, s>>>(pv, pmass, pstate, n, out);
    PF_CHECK_LAUNCH();
    return PF_OK;
}
