// Adaptive PF-FLIP step on an MSBG (pfflip/mr_sim.py:203-310,
// pfflip/mr_kernels.py, pfflip/particles.py).
//
// Faces follow the operator's convention: cell `a` owns its three low faces
// at face[a * 3 + axis].  Per-particle work (refinement map, P2G, G2P,
// resampling, advection) is one thread per particle; per-cell work is one
// thread per cell, fused so that each pass reads the neighbours only after
// the previous pass has finished everywhere:
//   finish        accumulators -> cell phase, liquid tag, face mass, u*
//   face_prepare  face beta, wall faces, the pre-force copy u*, gravity
//   rhs           integrated divergence (+ the optional density correction)
//   project       pressure gradient on owned faces, and du = u - u*
//   max_div       max |div u| after projection
// P2G adds into 64-bit fixed-point accumulators (one power-of-two scale per
// quantity, chosen on the host from the largest single contribution), so
// the sums do not depend on the order of the atomics.
#include "pf_msbg.hpp"

namespace pf {

int fill_u32(void* dst, int64_t n, uint32_t v, hipStream_t s);
int exclusive_scan(const int* in, int64_t n, int* out, int* tmp, hipStream_t s);
int64_t scan_tmp_ints(int64_t n);

constexpr int kAcc = 8;  // cmass, cwsum, fmass x3, fmom x3 (channel-major)

struct MRConst {
    float r0, rho_l, rho_g, rho_min, alpha_phi, rho_0l, rho_tilde_0, k_div;
    float gx, gy, gz, dt, nu_l, nu_g, alpha_cap;
    float lo[3], hi[3];
    int max_level;
    float mass_scale, mom_scale, w_scale;
    double inv_mass, inv_mom, inv_w;
};

static MRConst mrconst(const pf_mr_params& p) {
    MRConst c;
    c.r0 = p.r0;
    c.rho_l = p.rho_l;
    c.rho_g = p.rho_g;
    c.rho_min = p.rho_min;
    c.alpha_phi = p.alpha_phi;
    c.rho_0l = p.rho_0l;
    c.rho_tilde_0 = p.rho_tilde_0;
    c.k_div = p.k_div;
    c.gx = p.gravity[0];
    c.gy = p.gravity[1];
    c.gz = p.gravity[2];
    c.dt = p.dt;
    c.nu_l = p.nu_l;
    c.nu_g = p.nu_g;
    c.alpha_cap = p.alpha_cap;
    for (int a = 0; a < 3; ++a) {
        c.lo[a] = p.lo[a];
        c.hi[a] = p.hi[a];
    }
    c.max_level = p.max_level;
    c.mass_scale = (float)p.mass_scale;
    c.mom_scale = (float)p.mom_scale;
    c.w_scale = (float)p.w_scale;
    c.inv_mass = 1.0 / p.mass_scale;
    c.inv_mom = 1.0 / p.mom_scale;
    c.inv_w = 1.0 / p.w_scale;
    return c;
}

__device__ inline float cell_size(const MGrid& g, int l) { return g.h0 * (float)(1 << l); }

// eq. (7) with the level-0 constants (mr_kernels.py:141-146)
__device__ inline float phase_field(float rho, float rho_min, float alpha_phi, float rho_0l) {
    if (rho < rho_min) return 0.0f;
    return fminf(sqrtf(fmaxf(rho - rho_min, 0.0f) / (alpha_phi * rho_0l)), 1.0f);
}

// ------------------------------------------------------------------ 1. refinement map
// M[block] = min(l_p, M[block]) over the blocks a particle's box of half
// width cfl_eps touches (particles.py:49-85).  Neighbouring particles touch
// the same few blocks, so a wave combines its lanes first: each lane walks
// its blocks in index order; per round the first pending lane's block is
// taken, the lanes at that block reduce min(l), and one lane issues the
// atomic (the map only decreases, so it is skipped where the map already
// holds the value or less).
__global__ __launch_bounds__(256) void k_mr_refine_map(const float* __restrict__ x, const int* __restrict__ lev,
                                                       int64_t n, int* map, int nbx, int nby, int nbz, float bext,
                                                       float eps) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int i0 = 0, i1 = -1, j0 = 0, j1 = 0, k0 = 0, k1 = 0, l = 0;
    if (p < n) {
        const float px = x[3 * p], py = x[3 * p + 1], pz = x[3 * p + 2];
        l = lev[p];
        i0 = max((int)floorf((px - eps) / bext), 0);
        i1 = min((int)floorf((px + eps) / bext), nbx - 1);
        j0 = max((int)floorf((py - eps) / bext), 0);
        j1 = min((int)floorf((py + eps) / bext), nby - 1);
        k0 = max((int)floorf((pz - eps) / bext), 0);
        k1 = min((int)floorf((pz + eps) / bext), nbz - 1);
        if (j0 > j1 || k0 > k1) i1 = i0 - 1;  // box outside the raster
    }
    int i = i0, j = j0, k = k0;
    for (;;) {
        const bool pending = i <= i1;
        const unsigned long long act = __ballot(pending);
        if (act == 0ull) break;
        const int key = pending ? (i * nby + j) * nbz + k : -1;
        const int first = __ffsll((long long)act) - 1;
        const int lead = __shfl(key, first, 64);
        const bool mine = pending && key == lead;
        int m = mine ? l : 0x7fffffff;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
        if (lane == first && map[lead] > m) atomicMin(map + lead, m);
        if (mine) {
            if (++k > k1) {
                k = k0;
                if (++j > j1) {
                    j = j0;
                    ++i;
                }
            }
        }
    }
}

// max |v| over particles (mr_kernels.py:507-510)
__global__ __launch_bounds__(256) void k_mr_max_speed(const float* __restrict__ v, int64_t n, float* __restrict__ out) {
    float m = 0.0f;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const float a = v[3 * p], b = v[3 * p + 1], c = v[3 * p + 2];
        m = fmaxf(m, sqrtf(a * a + b * b + c * c));
    }
    __shared__ float sh[4];
    m = block_max(m, sh);
    if (threadIdx.x == 0) atomic_max_nonneg(out, m);
}

// ------------------------------------------------------------------ 4. P2G
// Mass (eq. 3) and momentum (eq. 4) on the low faces, mass and the bare
// weight sum at the centres, over the cells of every block the support
// sphere touches (mr_kernels.py:48-133).  m_p = rho 8^l_p, r_p = r0 2^l_p
// (particles.py:234-248).  `add(cell, values, mask)` receives the
// fixed-point contributions to one cell (bit ch of mask: channel ch set).
// round(v * scale) to the nearest integer (ties to even), for |v * scale| <
// 2^51: adding 1.5 * 2^52 in double leaves the rounded integer in the low
// mantissa bits (3 instructions; the float-to-int64 conversion takes 11).
// The host keeps every single contribution below 2^48.
__device__ inline long long fixed(float v, float scale) {
    const double d = (double)(v * scale) + 6755399441055744.0;
    return __double_as_longlong(d) - 0x4338000000000000LL;
}

template <class Add>
__device__ inline void splat_particle(const MGrid& g, const float* __restrict__ x, const float* __restrict__ v,
                                      const int* __restrict__ type, const int* __restrict__ lev, int64_t p,
                                      const MRConst& c, Add&& add) {
    const float px = x[3 * p], py = x[3 * p + 1], pz = x[3 * p + 2];
    const float vel[3] = {v[3 * p], v[3 * p + 1], v[3 * p + 2]};
    const int lp = lev[p];
    const float mp = (type[p] == 1 ? c.rho_l : c.rho_g) * (float)(1 << (3 * lp));
    const float rp = c.r0 * (float)(1 << lp);
    const float r2 = rp * rp, inv_r2 = 1.0f / r2;
    const float bext = (float)g.n0 * g.h0;
    const int bi0 = max((int)floorf((px - rp) / bext), 0), bi1 = min((int)floorf((px + rp) / bext), g.nbx - 1);
    const int bj0 = max((int)floorf((py - rp) / bext), 0), bj1 = min((int)floorf((py + rp) / bext), g.nby - 1);
    const int bk0 = max((int)floorf((pz - rp) / bext), 0), bk1 = min((int)floorf((pz + rp) / bext), g.nbz - 1);
    for (int bi = bi0; bi <= bi1; ++bi)
        for (int bj = bj0; bj <= bj1; ++bj)
            for (int bk = bk0; bk <= bk1; ++bk) {
                const int s = g.slot[(bi * g.nby + bj) * g.nbz + bk];
                if (s < 0) continue;
                const int l = g.blev[s];
                const int nn = g.n0 >> l;
                const float h = cell_size(g, l), half = 0.5f * h;
                const float ox = (float)bi * bext, oy = (float)bj * bext, oz = (float)bk * bext;
                // cells whose centre ((i + 1/2) h) or low face (i h) lies
                // within r: i in [(x - r)/h - 1/2, (x + r)/h].  A cell this
                // range misses by rounding has w < 2^-49 of a full weight,
                // which is 0 in fixed point (every scale keeps the largest
                // single contribution below 2^48), so the sums equal those
                // of the reference's wider walk (index range +-1).
                const int i0 = max((int)floorf((px - rp - ox) / h - 0.5f), 0);
                const int i1 = min((int)floorf((px + rp - ox) / h), nn - 1);
                const int j0 = max((int)floorf((py - rp - oy) / h - 0.5f), 0);
                const int j1 = min((int)floorf((py + rp - oy) / h), nn - 1);
                const int k0 = max((int)floorf((pz - rp - oz) / h - 0.5f), 0);
                const int k1 = min((int)floorf((pz + rp - oz) / h), nn - 1);
                const int base = g.coff[s];
                for (int i = i0; i <= i1; ++i) {
                    const float dxc = ox + ((float)i + 0.5f) * h - px, dxf = dxc - half;
                    const float xc2 = dxc * dxc, xf2 = dxf * dxf;
                    if (fminf(xc2, xf2) >= r2) continue;
                    for (int j = j0; j <= j1; ++j) {
                        const float dyc = oy + ((float)j + 0.5f) * h - py, dyf = dyc - half;
                        const float yc2 = dyc * dyc, yf2 = dyf * dyf;
                        if (fminf(xc2, xf2) + fminf(yc2, yf2) >= r2) continue;
                        for (int k = k0; k <= k1; ++k) {
                            const float dzc = oz + ((float)k + 0.5f) * h - pz, dzf = dzc - half;
                            const float zc2 = dzc * dzc, zf2 = dzf * dzf;
                            const int addr = base + (i * nn + j) * nn + k;
                            const float d2[4] = {xc2 + yc2 + zc2, xf2 + yc2 + zc2, xc2 + yf2 + zc2, xc2 + yc2 + zf2};
                            long long val[kAcc];
                            unsigned mask = 0u;
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                const float t = fmaxf(1.0f - d2[q] * inv_r2, 0.0f);
                                const float w = t * t * t;
                                if (!(w > 0.0f)) continue;
                                const float wm = w * mp;
                                if (q == 0) {
                                    val[0] = fixed(wm, c.mass_scale);
                                    val[1] = fixed(w, c.w_scale);
                                    mask |= 3u;
                                } else {
                                    val[1 + q] = fixed(wm, c.mass_scale);
                                    val[4 + q] = fixed(wm * vel[q - 1], c.mom_scale);
                                    mask |= (1u << (1 + q)) | (1u << (4 + q));
                                }
                            }
                            if (mask) add(addr, val, mask);
                        }
                    }
                }
            }
}

// one thread per particle in index order, every contribution a global atomic
__global__ __launch_bounds__(256) void k_mr_p2g(MGrid g, const float* __restrict__ x, const float* __restrict__ v,
                                                const int* __restrict__ type, const int* __restrict__ lev, int64_t n,
                                                MRConst c, unsigned long long* __restrict__ acc) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t nc = g.ncells;
    splat_particle(g, x, v, type, lev, p, c, [&](int addr, const long long* val, unsigned mask) {
#pragma unroll
        for (int ch = 0; ch < kAcc; ++ch)
            if (mask & (1u << ch)) atomicAdd(acc + ch * nc + addr, (unsigned long long)val[ch]);
    });
}

// Visit order for the cached P2G: particles sorted by the address of the
// cell they sit in (key ncells outside the grid); the order inside a cell
// is whatever the counting atomics give, which the fixed-point sums do not
// see.  The particle arrays themselves keep their order.
__global__ __launch_bounds__(256) void k_mr_cell_key(MGrid g, const float* __restrict__ x, int64_t n,
                                                     int* __restrict__ key, int* __restrict__ rank,
                                                     int* __restrict__ count) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const MCell cl = locate_at(g, x[3 * p], x[3 * p + 1], x[3 * p + 2]);
    const int k = cl.addr >= 0 ? cl.addr : g.ncells;
    key[p] = k;
    rank[p] = atomicAdd(count + k, 1);
}

__global__ __launch_bounds__(256) void k_mr_cell_scatter(const int* __restrict__ key, const int* __restrict__ rank,
                                                         const int* __restrict__ start, int64_t n,
                                                         int* __restrict__ perm) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    perm[start[key[p]] + rank[p]] = (int)p;
}

// P2G over the cell-sorted order.  A workgroup's particles are neighbours,
// so their contributions land on a few hundred cells: each workgroup keeps
// a hash of kP2GSlots cells in LDS (tag + the 8 fixed-point channels), adds
// there with LDS atomics, and flushes every cached cell with one global
// atomic per non-zero channel at the end.  A contribution whose slot holds
// another cell goes to global memory directly.  Integer sums, so the result
// equals the uncached kernel's bit for bit.
// 512 slots (34 KB of LDS) let four workgroups share a CU; 1024 slots with
// 512 particles per workgroup ran 6 % slower at two workgroups per CU
constexpr int kP2GSlots = 512;
constexpr int kP2GPerBlock = 256;  // particles per workgroup (1 per thread)
static_assert(kP2GSlots == 1 << 9, "the slot hash keeps 9 bits");

__global__ __launch_bounds__(256) void k_mr_p2g_cached(MGrid g, const float* __restrict__ x,
                                                       const float* __restrict__ v, const int* __restrict__ type,
                                                       const int* __restrict__ lev, const int* __restrict__ perm,
                                                       int64_t n, MRConst c, unsigned long long* __restrict__ acc) {
    __shared__ int tag[kP2GSlots];
    __shared__ unsigned long long lacc[kP2GSlots * kAcc];
    for (int s = threadIdx.x; s < kP2GSlots; s += blockDim.x) tag[s] = -1;
    for (int s = threadIdx.x; s < kP2GSlots * kAcc; s += blockDim.x) lacc[s] = 0ull;
    __syncthreads();
    const int64_t nc = g.ncells;
    const int64_t p0 = (int64_t)blockIdx.x * kP2GPerBlock;
    for (int q = threadIdx.x; q < kP2GPerBlock; q += blockDim.x) {
        if (p0 + q >= n) break;
        const int64_t p = perm[p0 + q];
        splat_particle(g, x, v, type, lev, p, c, [&](int addr, const long long* val, unsigned mask) {
            const int s = (int)(((unsigned)addr * 2654435761u) >> 23);  // 9-bit multiplicative hash
            int t = tag[s];
            if (t != addr && t == -1) {
                t = atomicCAS(tag + s, -1, addr);
                if (t == -1) t = addr;
            }
            if (t == addr) {
#pragma unroll
                for (int ch = 0; ch < kAcc; ++ch)
                    if (mask & (1u << ch)) atomicAdd(lacc + s * kAcc + ch, (unsigned long long)val[ch]);
            } else {
#pragma unroll
                for (int ch = 0; ch < kAcc; ++ch)
                    if (mask & (1u << ch)) atomicAdd(acc + ch * nc + addr, (unsigned long long)val[ch]);
            }
        });
    }
    __syncthreads();
    for (int s = threadIdx.x; s < kP2GSlots; s += blockDim.x) {
        const int a = tag[s];
        if (a < 0) continue;
#pragma unroll
        for (int ch = 0; ch < kAcc; ++ch) {
            const unsigned long long val = lacc[s * kAcc + ch];
            if (val != 0ull) atomicAdd(acc + ch * nc + a, val);
        }
    }
}

// ------------------------------------------------------------------ 5. cell finish
// cwsum; phase field and liquid tag with the per-level normalisation
// (mr_kernels.py:149-165); face mass and u* = P / m (mr_kernels.py:136-146).
// The accumulators are cleared for the next P2G.
__global__ __launch_bounds__(256) void k_mr_finish(MGrid g, MRConst c, int64_t* __restrict__ acc,
                                                   float* __restrict__ cwsum, float* __restrict__ phi,
                                                   int* __restrict__ is_liquid, float* __restrict__ fmass,
                                                   float* __restrict__ fvel) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    const int64_t nc = g.ncells;
    const int l = g.blev[g.cslot[a]];
    const float norm = (float)(1 << (3 * l));
    const float cm = (float)((double)acc[a] * c.inv_mass);
    cwsum[a] = (float)((double)acc[nc + a] * c.inv_w);
    const float f = phase_field(cm / norm, c.rho_min, c.alpha_phi, c.rho_0l);
    phi[a] = f;
    is_liquid[a] = f >= 0.5f ? 1 : 0;
    const float eps = 1e-12f * c.rho_0l;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const double m = (double)acc[(2 + ax) * nc + a] * c.inv_mass;
        const double mom = (double)acc[(5 + ax) * nc + a] * c.inv_mom;
        fmass[a * 3 + ax] = (float)m;
        fvel[a * 3 + ax] = (float)m > eps ? (float)(mom / m) : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < kAcc; ++q) acc[q * nc + a] = 0;
}

// ------------------------------------------------------------------ face coefficients, walls, forces
// beta = 1 / rho(phi_face) on owned faces, phi = 1 / 0 between two liquid /
// two air cells, beta = 0 on domain walls (mr_kernels.py:168-206); zero the
// wall velocities (mr_kernels.py:209-214); u* = u (mr_sim.py:238); gravity on
// open faces (mr_kernels.py:220-227)
__global__ __launch_bounds__(256) void k_mr_face_prepare(MGrid g, MRConst c, const int* __restrict__ is_liquid,
                                                         const float* __restrict__ fmass, float* __restrict__ fbeta,
                                                         float* __restrict__ fvel, float* __restrict__ fvel_star) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    BlockRef br;
    const int4 o = origin_ref(g, a, br);
    const float norm = (float)(1 << (3 * o.w));
    const int hi = is_liquid[a];
    const float gv[3] = {c.gx, c.gy, c.gz};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const MCell nb = locate_near(g, br, o.x - comp(ax, 0), o.y - comp(ax, 1), o.z - comp(ax, 2));
        float beta = 0.0f;
        if (nb.addr >= 0) {
            const int lo = is_liquid[nb.addr];
            float f;
            if (lo == 1 && hi == 1)
                f = 1.0f;
            else if (lo == 0 && hi == 0)
                f = 0.0f;
            else
                f = phase_field(fmass[a * 3 + ax] / norm, c.rho_min, c.alpha_phi, c.rho_0l);
            beta = 1.0f / ((1.0f - f) * c.rho_g + f * c.rho_l);
        }
        fbeta[a * 3 + ax] = beta;
        float u = beta == 0.0f ? 0.0f : fvel[a * 3 + ax];
        fvel_star[a * 3 + ax] = u;
        if (beta != 0.0f) u += gv[ax] * c.dt;
        fvel[a * 3 + ax] = u;
    }
}

// flux-weighted velocity through the high face: a finer neighbour backs the
// face with a 2x2 patch whose mean is taken (mr_kernels.py:230-283)
__device__ inline float face_velocity_high(const MGrid& g, const BlockRef& br, const float* __restrict__ fvel, int4 o,
                                          int ax) {
    const int ext = 1 << o.w;
    const int fx = o.x + comp(ax, 0) * ext, fy = o.y + comp(ax, 1) * ext, fz = o.z + comp(ax, 2) * ext;
    const MCell nb = locate_near(g, br, fx, fy, fz);
    if (nb.addr < 0) return 0.0f;
    if (nb.level >= o.w) return fvel[nb.addr * 3 + ax];
    const int sub = 1 << nb.level;
    const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
    float acc = 0.0f, cnt = 0.0f;
    for (int u = 0; u < 2; ++u)
        for (int v = 0; v < 2; ++v) {
            const int du = u * sub, dv = v * sub;
            const MCell nc = locate_near(g, br, fx + comp(a1, 0) * du + comp(a2, 0) * dv,
                                         fy + comp(a1, 1) * du + comp(a2, 1) * dv, fz + comp(a1, 2) * du + comp(a2, 2) * dv);
            if (nc.addr >= 0) {
                acc += fvel[nc.addr * 3 + ax];
                cnt += 1.0f;
            }
        }
    return cnt > 0.0f ? acc / cnt : 0.0f;
}

// b = -(1/dt) sum_faces u* . A_f, minus the optional density correction
// (mr_kernels.py:286-313)
__global__ __launch_bounds__(256) void k_mr_rhs(MGrid g, MRConst c, const float* __restrict__ fvel,
                                                const float* __restrict__ cwsum, float* __restrict__ rhs) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    BlockRef br;
    const int4 o = origin_ref(g, a, br);
    const float h = cell_size(g, o.w), area = h * h;
    float flux = 0.0f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) flux += (face_velocity_high(g, br, fvel, o, ax) - fvel[a * 3 + ax]) * area;
    float r = -flux / c.dt;
    if (c.k_div > 0.0f) {
        const float excess = (cwsum[a] / (float)(1 << (3 * o.w)) - c.rho_tilde_0) / c.rho_tilde_0;
        r -= c.k_div * fmaxf(excess, 0.0f) * h * h * h / c.dt;
    }
    rhs[a] = r;
}

// u = u* - dt beta (p_a - p_low) / dist on owned faces (mr_kernels.py:316-343),
// then du = u - u* (mr_kernels.py:415-418)
__global__ __launch_bounds__(256) void k_mr_project(MGrid g, MRConst c, const float* __restrict__ pres,
                                                    const float* __restrict__ fbeta, float* __restrict__ fvel,
                                                    const float* __restrict__ fvel_star, float* __restrict__ fdvel) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    BlockRef br;
    const int4 o = origin_ref(g, a, br);
    const float hA = cell_size(g, o.w);
    const float pa = pres[a];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        float u = fvel[a * 3 + ax];
        const float beta = fbeta[a * 3 + ax];
        if (beta != 0.0f) {
            const MCell nb = locate_near(g, br, o.x - comp(ax, 0), o.y - comp(ax, 1), o.z - comp(ax, 2));
            if (nb.addr >= 0) {
                const float dist = 0.5f * (hA + cell_size(g, nb.level));
                u -= c.dt * beta * (pa - pres[nb.addr]) / dist;
                fvel[a * 3 + ax] = u;
            }
        }
        fdvel[a * 3 + ax] = u - fvel_star[a * 3 + ax];
    }
}

// max |div u| (mr_kernels.py:346-353)
__global__ __launch_bounds__(256) void k_mr_max_div(MGrid g, const float* __restrict__ fvel, float* __restrict__ out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (a < g.ncells) {
        BlockRef br;
        const int4 o = origin_ref(g, a, br);
        float d = 0.0f;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) d += face_velocity_high(g, br, fvel, o, ax) - fvel[a * 3 + ax];
        m = fabsf(d / cell_size(g, o.w));
    }
    __shared__ float sh[4];
    m = block_max(m, sh);
    if (threadIdx.x == 0) atomic_max_nonneg(out, m);
}

// ------------------------------------------------------------------ G2P and advection
// trilinear face sample at the resolution of the cell containing x; taps
// resolve through locate, so a coarser or finer neighbour contributes its own
// face (mr_kernels.py:359-405)
// the cell containing a sample point and its block, resolved once for the
// three staggered components
struct SamplePoint {
    MCell c;
    BlockRef br;
};

__device__ inline SamplePoint sample_point(const MGrid& g, float x, float y, float z) {
    SamplePoint s;
    s.c = locate_at(g, x, y, z);
    if (s.c.addr >= 0) s.br = block_ref(g, s.c.slot);
    return s;
}

__device__ inline float sample_face(const MGrid& g, const float* __restrict__ f, float x, float y, float z, int ax,
                                    const SamplePoint& sp) {
    const MCell& c = sp.c;
    if (c.addr < 0) return 0.0f;
    const BlockRef& br = sp.br;
    const int ext = 1 << c.level;
    const float h = cell_size(g, c.level);
    const float sx = x / h - (ax == 0 ? 0.0f : 0.5f);
    const float sy = y / h - (ax == 1 ? 0.0f : 0.5f);
    const float sz = z / h - (ax == 2 ? 0.0f : 0.5f);
    const float fi = floorf(sx), fj = floorf(sy), fk = floorf(sz);
    const int i0 = (int)fi, j0 = (int)fj, k0 = (int)fk;
    const float tx = sx - fi, ty = sy - fj, tz = sz - fk;
    float acc = 0.0f, wsum = 0.0f;
    for (int di = 0; di < 2; ++di) {
        const float wi = di == 0 ? 1.0f - tx : tx;
        for (int dj = 0; dj < 2; ++dj) {
            const float wj = dj == 0 ? 1.0f - ty : ty;
            for (int dk = 0; dk < 2; ++dk) {
                const float wk = dk == 0 ? 1.0f - tz : tz;
                const float w = wi * wj * wk;
                if (w <= 0.0f) continue;
                const MCell nc = locate_near(g, br, (i0 + di) * ext, (j0 + dj) * ext, (k0 + dk) * ext);
                if (nc.addr < 0) continue;
                acc += w * f[nc.addr * 3 + ax];
                wsum += w;
            }
        }
    }
    return wsum > 0.0f ? acc / wsum : 0.0f;
}

// G2P samples two fields at one point through the same taps: the taps of
// the three staggered components (weight and cell address; -1: not summed,
// a zero weight or no cell) are resolved once and the 48 values loaded
// together (973 -> 830 us at 9.1M particles).  The advection's three
// samples at three points keep sample_face: with the taps batched per point
// it took 1.38 ms instead of 1.20 (more registers, the out-of-block lookups
// still serial).
struct Taps3 {
    float w[3][8];
    int addr[3][8];
};

__device__ inline void sample_taps(const MGrid& g, float x, float y, float z, const SamplePoint& sp, Taps3& t) {
    const MCell& c = sp.c;
    const BlockRef& br = sp.br;
    const int ext = 1 << c.level;
    const float h = cell_size(g, c.level);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float sx = x / h - (ax == 0 ? 0.0f : 0.5f);
        const float sy = y / h - (ax == 1 ? 0.0f : 0.5f);
        const float sz = z / h - (ax == 2 ? 0.0f : 0.5f);
        const float fi = floorf(sx), fj = floorf(sy), fk = floorf(sz);
        const int i0 = (int)fi, j0 = (int)fj, k0 = (int)fk;
        const float tx = sx - fi, ty = sy - fj, tz = sz - fk;
#pragma unroll
        for (int di = 0; di < 2; ++di) {
            const float wi = di == 0 ? 1.0f - tx : tx;
#pragma unroll
            for (int dj = 0; dj < 2; ++dj) {
                const float wj = dj == 0 ? 1.0f - ty : ty;
#pragma unroll
                for (int dk = 0; dk < 2; ++dk) {
                    const float wk = dk == 0 ? 1.0f - tz : tz;
                    const int q = di * 4 + dj * 2 + dk;
                    const float w = wi * wj * wk;
                    t.w[ax][q] = w;
                    t.addr[ax][q] = w > 0.0f
                                        ? locate_near(g, br, (i0 + di) * ext, (j0 + dj) * ext, (k0 + dk) * ext).addr
                                        : -1;
                }
            }
        }
    }
}

// The tap values of field f (address -1 reads entry 0, masked in the sum):
// all 24 loads are issued before the first is used.
__device__ inline void tap_values(const float* __restrict__ f, const Taps3& t, float (&v)[3][8]) {
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int q = 0; q < 8; ++q) v[ax][q] = f[max(t.addr[ax][q], 0) * 3 + ax];
}

// trilinear face samples from the taps, normalised by the weight of the taps
// that exist, summed in tap order (mr_kernels.py:359-405)
__device__ inline void tap_sum(const SamplePoint& sp, const Taps3& t, const float (&v)[3][8], float (&out)[3]) {
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        float acc = 0.0f, wsum = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (t.addr[ax][q] >= 0) {
                acc += t.w[ax][q] * v[ax][q];
                wsum += t.w[ax][q];
            }
        out[ax] = sp.c.addr < 0 ? 0.0f : (wsum > 0.0f ? acc / wsum : 0.0f);
    }
}

// alpha_FLIP from the local grid spacing (particles.py:251-276), then
// u_p = alpha (u_p + I(du)) + (1 - alpha) I(u), eq. (12) (mr_kernels.py:413-430)
__global__ __launch_bounds__(256) void k_mr_g2p(MGrid g, const float* __restrict__ x, float* __restrict__ v,
                                                const int* __restrict__ type, int64_t n, MRConst c,
                                                const float* __restrict__ fvel, const float* __restrict__ fdvel) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float px = x[3 * p], py = x[3 * p + 1], pz = x[3 * p + 2];
    const SamplePoint sp = sample_point(g, px, py, pz);
    const MCell& cl = sp.c;
    const int lv = cl.addr >= 0 ? cl.level : 0;
    const float dx = cell_size(g, lv);
    const float nu = type[p] == 1 ? c.nu_l : c.nu_g;
    const float alpha = fminf(fmaxf(1.0f - 6.0f * nu * c.dt / (dx * dx), 0.0f), c.alpha_cap);
    // both fields through the same taps, all 48 values loaded at once
    float delta[3] = {0.0f, 0.0f, 0.0f}, pic[3] = {0.0f, 0.0f, 0.0f};
    if (cl.addr >= 0) {
        Taps3 t;
        sample_taps(g, px, py, pz, sp, t);
        float vd[3][8], vu[3][8];
        tap_values(fdvel, t, vd);
        tap_values(fvel, t, vu);
        tap_sum(sp, t, vd, delta);
        tap_sum(sp, t, vu, pic);
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) v[3 * p + ax] = alpha * (v[3 * p + ax] + delta[ax]) + (1.0f - alpha) * pic[ax];
}

// Ralston RK3 through the MSBG velocity, clamped to [lo, hi]
// (mr_kernels.py:421-438), then the wall reflection of the velocity
// (mr_kernels.py:441-461)
__global__ __launch_bounds__(256) void k_mr_advect(MGrid g, float* __restrict__ x, float* __restrict__ v, int64_t n,
                                                   MRConst c, const float* __restrict__ fvel) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float q[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
    float k1[3], k2[3], k3[3], t[3];
    SamplePoint sp = sample_point(g, q[0], q[1], q[2]);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) k1[ax] = sample_face(g, fvel, q[0], q[1], q[2], ax, sp);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) t[ax] = q[ax] + 0.5f * c.dt * k1[ax];
    sp = sample_point(g, t[0], t[1], t[2]);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) k2[ax] = sample_face(g, fvel, t[0], t[1], t[2], ax, sp);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) t[ax] = q[ax] + 0.75f * c.dt * k2[ax];
    sp = sample_point(g, t[0], t[1], t[2]);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) k3[ax] = sample_face(g, fvel, t[0], t[1], t[2], ax, sp);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        float xa = q[ax] + c.dt * ((2.0f / 9.0f) * k1[ax] + (3.0f / 9.0f) * k2[ax] + (4.0f / 9.0f) * k3[ax]);
        xa = fminf(fmaxf(xa, c.lo[ax]), c.hi[ax]);
        float va = v[3 * p + ax];
        if (xa <= c.lo[ax])
            va = fmaxf(va, 0.0f);
        else if (xa >= c.hi[ax])
            va = fminf(va, 0.0f);
        x[3 * p + ax] = xa;
        v[3 * p + ax] = va;
    }
}

// ------------------------------------------------------------------ 6. resampling
// A block is "interface" when its cells hold both phases (mr_kernels.py:464-477);
// one workgroup per active block reduces its contiguous cell range.
__global__ __launch_bounds__(256) void k_mr_block_phase(MGrid g, const int* __restrict__ is_liquid,
                                                        int* __restrict__ mixed) {
    const int s = blockIdx.x;
    const int c0 = g.coff[s];
    const int nn = g.n0 >> g.blev[s];
    const int cnt = nn * nn * nn;
    int liq = 0, air = 0;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const int t = is_liquid[c0 + i];
        liq |= t;
        air |= 1 - t;
    }
    __shared__ int sh[2][4];
    liq = __any(liq) ? 1 : 0;
    air = __any(air) ? 1 : 0;
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = liq;
        sh[1][threadIdx.x >> 6] = air;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int l = 0, a = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            l |= sh[0][w];
            a |= sh[1][w];
        }
        mixed[g.blk[s]] = (l && a) ? 1 : 0;
    }
}

constexpr int kInfD2 = 1 << 30;

// exact Euclidean distance to the nearest interface block, in block units,
// as separable squared-distance passes along z, y, x (the reference calls
// scipy.ndimage.distance_transform_edt: particles.py:387-419)
__global__ __launch_bounds__(256) void k_mr_edt(int pass, const int* __restrict__ mixed, const int* __restrict__ in,
                                                int* __restrict__ out, float* __restrict__ dist, int nbx, int nby,
                                                int nbz, float max_dist) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbx * nby * nbz) return;
    const int k = t % nbz, j = (t / nbz) % nby, i = t / (nbz * nby);
    int best = kInfD2;
    if (pass == 0) {
        for (int kk = 0; kk < nbz; ++kk)
            if (mixed[(i * nby + j) * nbz + kk]) best = min(best, (k - kk) * (k - kk));
        out[t] = best;
    } else if (pass == 1) {
        for (int jj = 0; jj < nby; ++jj) {
            const int g = in[(i * nby + jj) * nbz + k];
            if (g < kInfD2) best = min(best, g + (j - jj) * (j - jj));
        }
        out[t] = best;
    } else {
        for (int ii = 0; ii < nbx; ++ii) {
            const int g = in[(ii * nby + j) * nbz + k];
            if (g < kInfD2) best = min(best, g + (i - ii) * (i - ii));
        }
        dist[t] = best >= kInfD2 ? max_dist : fminf((float)sqrt((double)best), max_dist);
    }
}

// l_new = floor(|beta_bw d~|) with d~ interpolated trilinearly from the block
// raster, clamped to [0, max_level] and to one level per step; deferred
// probabilistic coarsening (particles.py:88-168).  nout = 0, 1 or 8.
__global__ __launch_bounds__(256) void k_mr_classify(const float* __restrict__ x, const int* __restrict__ lev,
                                                     int* __restrict__ defer, const float* __restrict__ dist, int nbx,
                                                     int nby, int nbz, float bext, float beta_bw, int max_level,
                                                     int threshold, int seed, int* __restrict__ lnew_out,
                                                     int* __restrict__ nout, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float gx = x[3 * p] / bext - 0.5f, gy = x[3 * p + 1] / bext - 0.5f, gz = x[3 * p + 2] / bext - 0.5f;
    const float fi = floorf(gx), fj = floorf(gy), fk = floorf(gz);
    const int i0 = (int)fi, j0 = (int)fj, k0 = (int)fk;
    const float tx = gx - fi, ty = gy - fj, tz = gz - fk;
    float d = 0.0f;
    for (int a = 0; a < 2; ++a) {
        const float wa = a == 0 ? 1.0f - tx : tx;
        const int ii = min(max(i0 + a, 0), nbx - 1);
        for (int b = 0; b < 2; ++b) {
            const float wb = b == 0 ? 1.0f - ty : ty;
            const int jj = min(max(j0 + b, 0), nby - 1);
            for (int cc = 0; cc < 2; ++cc) {
                const float wc = cc == 0 ? 1.0f - tz : tz;
                const int kk = min(max(k0 + cc, 0), nbz - 1);
                d += wa * wb * wc * dist[(ii * nby + jj) * nbz + kk];
            }
        }
    }
    const int lp = lev[p];
    int ln = min(max((int)floorf(fabsf(beta_bw * d)), 0), max_level);
    ln = min(max(ln, lp - 1), lp + 1);
    lnew_out[p] = ln;
    if (ln < lp) {
        defer[p] = 0;
        nout[p] = 8;
        return;
    }
    if (ln > lp) {
        const int dp = defer[p] + 1;
        defer[p] = dp;
        if (dp >= threshold) {
            uint32_t st = rng_init(seed, (int32_t)p);
            const float prob_del = 1.0f - ldexpf(1.0f, -3 * (ln - lp));
            if (rng_uniform(st) < prob_del) {
                nout[p] = 0;
                return;
            }
        }
        nout[p] = 1;
        return;
    }
    defer[p] = 0;
    nout[p] = 1;
}

// survivors (promoted when the counter fired) and 8 split children jittered
// inside the octants of the parent's cell (particles.py:171-231)
__global__ __launch_bounds__(256) void k_mr_emit(const float* __restrict__ x, const float* __restrict__ v,
                                                 const int* __restrict__ type, const int* __restrict__ lev,
                                                 const int* __restrict__ defer, const int* __restrict__ lnew,
                                                 const int* __restrict__ nout, const int* __restrict__ off, int64_t n,
                                                 int64_t cap, int threshold, float r0, int seed, float* __restrict__ ox,
                                                 float* __restrict__ ov, int* __restrict__ otype,
                                                 int* __restrict__ olev, int* __restrict__ odefer) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int cnt = nout[p];
    const int64_t o = off[p];
    if (cnt == 0 || o + cnt > cap) return;  // the host sized the output from offsets[n]
    const float px = x[3 * p], py = x[3 * p + 1], pz = x[3 * p + 2];
    const float vx = v[3 * p], vy = v[3 * p + 1], vz = v[3 * p + 2];
    const int t = type[p], lp = lev[p], ln = lnew[p];
    if (cnt == 1) {
        ox[3 * o] = px;
        ox[3 * o + 1] = py;
        ox[3 * o + 2] = pz;
        ov[3 * o] = vx;
        ov[3 * o + 1] = vy;
        ov[3 * o + 2] = vz;
        otype[o] = t;
        const bool promote = ln > lp && defer[p] >= threshold;
        olev[o] = promote ? ln : lp;
        odefer[o] = promote ? 0 : defer[p];
        return;
    }
    const float hp = r0 * (float)(1 << lp);
    const float half = 0.5f * hp;
    uint32_t st = rng_init(seed, (int32_t)p);
    for (int ch = 0; ch < 8; ++ch) {
        const float sx = (float)(ch & 1) - 0.5f, sy = (float)((ch >> 1) & 1) - 0.5f, sz = (float)((ch >> 2) & 1) - 0.5f;
        const float jx = (rng_uniform(st) - 0.5f) * half;
        const float jy = (rng_uniform(st) - 0.5f) * half;
        const float jz = (rng_uniform(st) - 0.5f) * half;
        const int64_t q = o + ch;
        ox[3 * q] = px + sx * half + jx;
        ox[3 * q + 1] = py + sy * half + jy;
        ox[3 * q + 2] = pz + sz * half + jz;
        ov[3 * q] = vx;
        ov[3 * q + 1] = vy;
        ov[3 * q + 2] = vz;
        otype[q] = t;
        olev[q] = lp - 1;
        odefer[q] = 0;
    }
}

}  // namespace pf

using namespace pf;

static int check_mgrid(const pf_msbg_grid* g, const char* who) {
    PF_REQUIRE(g && g->ncells > 0 && g->nslots > 0 && g->n0 == (1 << g->sh0) && g->h0 > 0.f, "%s: bad grid", who);
    PF_REQUIRE(g->level && g->slot && g->blk && g->blev && g->coff && g->cslot, "%s: null grid array", who);
    // face fields are indexed 3 * cell + axis in 32-bit arithmetic
    PF_REQUIRE(g->ncells <= kMsbgMaxCells, "%s: %d cells exceed the limit of %d", who, g->ncells, kMsbgMaxCells);
    return PF_OK;
}

static int check_parts(const pf_mr_particles* ps, const char* who) {
    PF_REQUIRE(ps && ps->n >= 0 && ps->n < ((int64_t)1 << 31), "%s: bad particle set", who);
    PF_REQUIRE(ps->n == 0 || (ps->x && ps->v && ps->type && ps->level && ps->defer), "%s: null particle array",
               who);
    return PF_OK;
}

static unsigned pgrid(int64_t n) { return cdiv(n > 0 ? n : 1, 256); }

PF_API int pf_mr_refine_map(const pf_mr_particles* ps, int32_t* level_map, int32_t nbx, int32_t nby, int32_t nbz,
                            float block_extent, float cfl_eps, int32_t max_level, pf_stream_t stream) {
    int rc = check_parts(ps, "pf_mr_refine_map");
    if (rc) return rc;
    PF_REQUIRE(level_map && nbx > 0 && nby > 0 && nbz > 0 && block_extent > 0.f && max_level >= 0,
               "pf_mr_refine_map: bad block raster");
    rc = fill_u32(level_map, (int64_t)nbx * nby * nbz, (uint32_t)max_level, S(stream));
    if (rc) return rc;
    if (ps->n == 0) return PF_OK;
    k_mr_refine_map<<<pgrid(ps->n), 256, 0, S(stream)>>>(ps->x, ps->level, ps->n, level_map, nbx, nby, nbz,
                                                         block_extent, cfl_eps);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_max_speed(const float* v, int64_t n, float* out, pf_stream_t stream) {
    PF_REQUIRE(out && (n == 0 || v), "pf_mr_max_speed: null pointer");
    int rc = fill_u32(out, 1, 0u, S(stream));
    if (rc || n == 0) return rc;
    const unsigned nb = cdiv(n, 256) < 2048u ? cdiv(n, 256) : 2048u;
    k_mr_max_speed<<<nb, 256, 0, S(stream)>>>(v, n, out);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_p2g(const pf_msbg_grid* g, const pf_mr_particles* ps, const pf_mr_params* prm, pf_mr_fields* f,
                     pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_p2g");
    if (rc) return rc;
    rc = check_parts(ps, "pf_mr_p2g");
    if (rc) return rc;
    PF_REQUIRE(prm && f && f->acc, "pf_mr_p2g: null params / accumulators");
    PF_REQUIRE(prm->mass_scale > 0.0 && prm->mom_scale > 0.0 && prm->w_scale > 0.0, "pf_mr_p2g: bad scales");
    if (ps->n == 0) return PF_OK;
    k_mr_p2g<<<pgrid(ps->n), 256, 0, S(stream)>>>(mgrid(*g), ps->x, ps->v, ps->type, ps->level, ps->n, mrconst(*prm),
                                                  reinterpret_cast<unsigned long long*>(f->acc));
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int64_t pf_mr_p2g_scratch_ints(int64_t n, int32_t ncells) {
    // key, rank, perm (n each); count (ncells + 1); start (ncells + 2); scan
    return 3 * (n > 0 ? n : 1) + 2 * ((int64_t)ncells + 2) + scan_tmp_ints((int64_t)ncells + 1);
}

PF_API int pf_mr_p2g_sorted(const pf_msbg_grid* g, const pf_mr_particles* ps, const pf_mr_params* prm,
                            pf_mr_fields* f, int32_t* scratch, pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_p2g_sorted");
    if (rc) return rc;
    rc = check_parts(ps, "pf_mr_p2g_sorted");
    if (rc) return rc;
    PF_REQUIRE(prm && f && f->acc && scratch, "pf_mr_p2g_sorted: null params / accumulators / scratch");
    PF_REQUIRE(prm->mass_scale > 0.0 && prm->mom_scale > 0.0 && prm->w_scale > 0.0, "pf_mr_p2g_sorted: bad scales");
    const int64_t n = ps->n;
    if (n == 0) return PF_OK;
    const int64_t nk = (int64_t)g->ncells + 1;
    int* key = scratch;
    int* rank = key + n;
    int* perm = rank + n;
    int* count = perm + n;
    int* start = count + (nk + 1);
    int* tmp = start + (nk + 1);
    hipStream_t s = S(stream);
    rc = fill_u32(count, nk, 0u, s);
    if (rc) return rc;
    const MGrid m = mgrid(*g);
    k_mr_cell_key<<<pgrid(n), 256, 0, s>>>(m, ps->x, n, key, rank, count);
    PF_CHECK_LAUNCH();
    rc = exclusive_scan(count, nk, start, tmp, s);
    if (rc) return rc;
    k_mr_cell_scatter<<<pgrid(n), 256, 0, s>>>(key, rank, start, n, perm);
    PF_CHECK_LAUNCH();
    k_mr_p2g_cached<<<cdiv(n, kP2GPerBlock), 256, 0, s>>>(m, ps->x, ps->v, ps->type, ps->level, perm, n,
                                                           mrconst(*prm), reinterpret_cast<unsigned long long*>(f->acc));
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_finish(const pf_msbg_grid* g, const pf_mr_params* prm, pf_mr_fields* f, pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_finish");
    if (rc) return rc;
    PF_REQUIRE(prm && f && f->acc && f->cwsum && f->phi && f->is_liquid && f->fmass && f->fvel,
               "pf_mr_finish: null field");
    k_mr_finish<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), mrconst(*prm), f->acc, f->cwsum, f->phi,
                                                             f->is_liquid, f->fmass, f->fvel);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_face_prepare(const pf_msbg_grid* g, const pf_mr_params* prm, pf_mr_fields* f, pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_face_prepare");
    if (rc) return rc;
    PF_REQUIRE(prm && f && f->is_liquid && f->fmass && f->fbeta && f->fvel && f->fvel_star,
               "pf_mr_face_prepare: null field");
    k_mr_face_prepare<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), mrconst(*prm), f->is_liquid, f->fmass,
                                                                   f->fbeta, f->fvel, f->fvel_star);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_rhs(const pf_msbg_grid* g, const pf_mr_params* prm, pf_mr_fields* f, pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_rhs");
    if (rc) return rc;
    PF_REQUIRE(prm && f && f->fvel && f->cwsum && f->rhs, "pf_mr_rhs: null field");
    PF_REQUIRE(prm->dt > 0.f, "pf_mr_rhs: dt must be positive");
    k_mr_rhs<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), mrconst(*prm), f->fvel, f->cwsum, f->rhs);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_project(const pf_msbg_grid* g, const pf_mr_params* prm, pf_mr_fields* f, pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_project");
    if (rc) return rc;
    PF_REQUIRE(prm && f && f->pressure && f->fbeta && f->fvel && f->fvel_star && f->fdvel,
               "pf_mr_project: null field");
    k_mr_project<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), mrconst(*prm), f->pressure, f->fbeta,
                                                              f->fvel, f->fvel_star, f->fdvel);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_max_divergence(const pf_msbg_grid* g, const float* fvel, float* out, pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_max_divergence");
    if (rc) return rc;
    PF_REQUIRE(fvel && out, "pf_mr_max_divergence: null pointer");
    rc = fill_u32(out, 1, 0u, S(stream));
    if (rc) return rc;
    k_mr_max_div<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), fvel, out);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_g2p(const pf_msbg_grid* g, const pf_mr_particles* ps, const pf_mr_params* prm,
                     const pf_mr_fields* f, pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_g2p");
    if (rc) return rc;
    rc = check_parts(ps, "pf_mr_g2p");
    if (rc) return rc;
    PF_REQUIRE(prm && f && f->fvel && f->fdvel, "pf_mr_g2p: null field");
    if (ps->n == 0) return PF_OK;
    k_mr_g2p<<<pgrid(ps->n), 256, 0, S(stream)>>>(mgrid(*g), ps->x, ps->v, ps->type, ps->n, mrconst(*prm), f->fvel,
                                                  f->fdvel);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_advect(const pf_msbg_grid* g, const pf_mr_particles* ps, const pf_mr_params* prm,
                        const float* fvel, pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_advect");
    if (rc) return rc;
    rc = check_parts(ps, "pf_mr_advect");
    if (rc) return rc;
    PF_REQUIRE(prm && fvel, "pf_mr_advect: null pointer");
    if (ps->n == 0) return PF_OK;
    k_mr_advect<<<pgrid(ps->n), 256, 0, S(stream)>>>(mgrid(*g), ps->x, ps->v, ps->n, mrconst(*prm), fvel);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_mr_block_distance(const pf_msbg_grid* g, const int32_t* is_liquid, int32_t* mixed, int32_t* tmp,
                                float* dist, float max_dist, pf_stream_t stream) {
    int rc = check_mgrid(g, "pf_mr_block_distance");
    if (rc) return rc;
    PF_REQUIRE(is_liquid && mixed && tmp && dist, "pf_mr_block_distance: null pointer");
    const int nblk = g->nbx * g->nby * g->nbz;
    rc = fill_u32(mixed, nblk, 0u, S(stream));
    if (rc) return rc;
    k_mr_block_phase<<<g->nslots, 256, 0, S(stream)>>>(mgrid(*g), is_liquid, mixed);
    PF_CHECK_LAUNCH();
    int* a = tmp;
    int* b = tmp + nblk;
    k_mr_edt<<<cdiv(nblk, 256), 256, 0, S(stream)>>>(0, mixed, nullptr, a, nullptr, g->nbx, g->nby, g->nbz, max_dist);
    PF_CHECK_LAUNCH();
    k_mr_edt<<<cdiv(nblk, 256), 256, 0, S(stream)>>>(1, mixed, a, b, nullptr, g->nbx, g->nby, g->nbz, max_dist);
    PF_CHECK_LAUNCH();
    k_mr_edt<<<cdiv(nblk, 256), 256, 0, S(stream)>>>(2, mixed, b, nullptr, dist, g->nbx, g->nby, g->nbz, max_dist);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int64_t pf_mr_scan_tmp_ints(int64_t n) { return scan_tmp_ints(n > 0 ? n : 1); }

PF_API int pf_mr_classify(const pf_mr_particles* ps, const float* dist, int32_t nbx, int32_t nby, int32_t nbz,
                          float block_extent, float beta_bw, int32_t max_level, int32_t defer_threshold, int32_t seed,
                          int32_t* lnew, int32_t* nout, int32_t* offsets, int32_t* scan_tmp, pf_stream_t stream) {
    int rc = check_parts(ps, "pf_mr_classify");
    if (rc) return rc;
    PF_REQUIRE(dist && lnew && nout && offsets && scan_tmp && nbx > 0 && nby > 0 && nbz > 0 && block_extent > 0.f,
               "pf_mr_classify: bad arguments");
    if (ps->n == 0) return fill_u32(offsets, 1, 0u, S(stream));
    k_mr_classify<<<pgrid(ps->n), 256, 0, S(stream)>>>(ps->x, ps->level, ps->defer, dist, nbx, nby, nbz, block_extent,
                                                       beta_bw, max_level, defer_threshold, seed, lnew, nout, ps->n);
    PF_CHECK_LAUNCH();
    return exclusive_scan(nout, ps->n, offsets, scan_tmp, S(stream));
}

PF_API int pf_mr_emit(const pf_mr_particles* in, const int32_t* lnew, const int32_t* nout, const int32_t* offsets,
                      int32_t defer_threshold, float r0, int32_t seed, pf_mr_particles* out, pf_stream_t stream) {
    int rc = check_parts(in, "pf_mr_emit");
    if (rc) return rc;
    rc = check_parts(out, "pf_mr_emit");
    if (rc) return rc;
    PF_REQUIRE(lnew && nout && offsets, "pf_mr_emit: null pointer");
    if (in->n == 0) return PF_OK;
    k_mr_emit<<<pgrid(in->n), 256, 0, S(stream)>>>(in->x, in->v, in->type, in->level, in->defer, lnew, nout, offsets,
                                                   in->n, out->n, defer_threshold, r0, seed, out->x, out->v, out->type,
                                                   out->level, out->defer);
    PF_CHECK_LAUNCH();
    return PF_OK;
}
