// Escaped particles, Lagrangian droplets and spray (pfflip/droplets.py,
// pfflip/sim.py:343-401).
//
// The reference launches classify_escaped, emit_spray_in_shear and
// advance_droplets one after the other over all particles.  Each of them
// reads and writes only its own particle (plus grid fields that none of them
// writes), so one pass that runs the three in order per particle gives the
// same result with a single read and write of the particle arrays.
//
// Random numbers: a counter-based stream per (seed, particle), built on the
// PCG hash of Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 9(3)
// 2020 (the RXS-M-XS output of O'Neill's PCG), keyed the way the reference
// keys Warp's rand_init(seed, p) (droplets.py:91,149).  Uniforms take the top
// 24 bits; normals are Box-Muller.
#include <hip/hip_fp16.h>

#include "pf_common.hpp"

namespace pf {

// ------------------------------------------------------------------ sampling
// Trilinear interpolation of a cell-centred field, indices clamped per axis
// (droplets.py:42-61).
__device__ inline float sample_cell(const float* __restrict__ f, float x, float y, float z, float dx,
                                    const Dims& d) {
    const float gx = x / dx - 0.5f, gy = y / dx - 0.5f, gz = z / dx - 0.5f;
    const int i0 = (int)floorf(gx), j0 = (int)floorf(gy), k0 = (int)floorf(gz);
    const float tx = gx - (float)i0, ty = gy - (float)j0, tz = gz - (float)k0;
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const float wa = a == 0 ? 1.0f - tx : tx;
        const int ii = min(max(i0 + a, 0), d.nx - 1);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const float wb = b == 0 ? 1.0f - ty : ty;
            const int jj = min(max(j0 + b, 0), d.ny - 1);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const float wc = c == 0 ? 1.0f - tz : tz;
                const int kk = min(max(k0 + c, 0), d.nz - 1);
                acc += wa * wb * wc * f[cell_index(d, ii, jj, kk)];
            }
        }
    }
    return acc;
}

// Face field of axis A at x (droplets.py:281-302; same taps as G2P).
template <int A>
__device__ inline float sample_face_at(const float* __restrict__ f, float x, float y, float z, float dx,
                                       const Dims& d) {
    const int n0 = d.nx + (A == 0), n1 = d.ny + (A == 1), n2 = d.nz + (A == 2);
    const float gx = x / dx - (A == 0 ? 0.f : 0.5f);
    const float gy = y / dx - (A == 1 ? 0.f : 0.5f);
    const float gz = z / dx - (A == 2 ? 0.f : 0.5f);
    const int i0 = (int)floorf(gx), j0 = (int)floorf(gy), k0 = (int)floorf(gz);
    const float tx = gx - (float)i0, ty = gy - (float)j0, tz = gz - (float)k0;
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const float wa = a == 0 ? 1.0f - tx : tx;
        const int ii = min(max(i0 + a, 0), n0 - 1);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const float wb = b == 0 ? 1.0f - ty : ty;
            const int jj = min(max(j0 + b, 0), n1 - 1);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const float wc = c == 0 ? 1.0f - tz : tz;
                const int kk = min(max(k0 + c, 0), n2 - 1);
                acc += wa * wb * wc * f[((int64_t)ii * n1 + jj) * n2 + kk];
            }
        }
    }
    return acc;
}

// One count per lane that passes `c`, one atomic per wave.
__device__ inline void wave_count(int32_t* ctr, bool c) {
    const uint64_t m = __ballot(c);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    if (lane == __ffsll((unsigned long long)m) - 1) atomicAdd(ctr, (int32_t)__popcll(m));
}

struct DropArgs {
    float* px;
    float* pv;
    const int32_t* ptype;
    int32_t* pstate;
    float* pdrop_r;
    int64_t n;
    Dims d;
    int32_t seed, max_substeps;
    const float* phi;
    const float* shear;
    const float* vel[3];
    float dx, dt, delta_phi, r_mu, r_sigma, r_min, r_max, spray_threshold, spray_rate;
    float g[3];
    float rho_g, rho_l, cd, evap_k;
    float lo[3], hi[3];
    int32_t* counters;
};

constexpr int kFlip = 0, kDroplet = 1, kDead = 2, kLiquid = 1;

__global__ __launch_bounds__(256) void k_droplets(DropArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = p < a.n;
    int st = in ? a.pstate[p] : kDead;
    bool emitted = false, bubble = false, merged = false, spray = false, alive = false;
    if (st != kDead) {
        const Dims d = a.d;
        float x = a.px[3 * p], y = a.px[3 * p + 1], z = a.px[3 * p + 2];
        const int pt = a.ptype[p];
        const int st0 = st;
        float r = a.pdrop_r[p];
        bool r_new = false;
        float r_emit = 0.f;

        // classify_escaped (droplets.py:67-106)
        const float f = sample_cell(a.phi, x, y, z, a.dx, d);
        if (st == kFlip) {
            if (pt == kLiquid && f < 0.5f - a.delta_phi) {
                st = kDroplet;
                uint32_t rs = rng_init(a.seed, (int32_t)p);
                r = fminf(fmaxf(expf(a.r_mu + a.r_sigma * rng_normal(rs)), a.r_min), a.r_max);
                r_new = true;
                r_emit = r;
                emitted = true;
            } else if (pt != kLiquid && f > 0.5f + a.delta_phi) {
                st = kDead;
                bubble = true;
            }
        } else if (st == kDroplet) {
            if (f > 0.5f + a.delta_phi) {
                st = kFlip;
                merged = true;
            }
        }

        // emit_spray_in_shear (droplets.py:112-153)
        if (a.shear && st == kFlip && pt == kLiquid && !(f > 0.9f || f < 0.1f) &&
            !(sample_cell(a.shear, x, y, z, a.dx, d) < a.spray_threshold)) {
            uint32_t rs = rng_init(a.seed + 104729, (int32_t)p);
            if (!(rng_uniform(rs) > a.spray_rate)) {
                st = kDroplet;
                r = fminf(fmaxf(expf(a.r_mu + a.r_sigma * rng_normal(rs)), a.r_min), a.r_max);
                r_new = true;
                r_emit = r;
                spray = true;
            }
        }

        // advance_droplets (droplets.py:211-278)
        bool moved = false;
        float vx = 0.f, vy = 0.f, vz = 0.f;
        if (st == kDroplet) {
            vx = a.pv[3 * p];
            vy = a.pv[3 * p + 1];
            vz = a.pv[3 * p + 2];
            const float speed = sqrtf(vx * vx + vy * vy + vz * vz) + 1e-8f;
            // the float is capped before the conversion; any cap above
            // max_substeps gives the same count
            const float cfl = fminf(speed * a.dt / (0.5f * a.dx), 1.0e9f);
            const int nsub = min(max((int)cfl + 1, 1), a.max_substeps);
            const float h = a.dt / (float)nsub;
            for (int s = 0; s < nsub && st == kDroplet; ++s) {
                const float ux = sample_face_at<0>(a.vel[0], x, y, z, a.dx, d);
                const float uy = sample_face_at<1>(a.vel[1], x, y, z, a.dx, d);
                const float uz = sample_face_at<2>(a.vel[2], x, y, z, a.dx, d);
                // sphere drag per unit mass, 3 rho_g Cd |w| w / (8 rho_l r)
                const float wx = ux - vx, wy = uy - vy, wz = uz - vz;
                const float sw = sqrtf(wx * wx + wy * wy + wz * wz);
                float c = 0.f;
                if (!(sw <= 0.f || r <= 0.f)) c = 3.0f * a.rho_g * a.cd * sw / (8.0f * a.rho_l * r);
                vx = vx + h * (a.g[0] + c * wx);
                vy = vy + h * (a.g[1] + c * wy);
                vz = vz + h * (a.g[2] + c * wz);
                x = x + h * vx;
                y = y + h * vy;
                z = z + h * vz;
                // evaporation dr/dt = -k/r, midpoint RK2
                if (a.evap_k > 0.f && r > 0.f) {
                    const float k1 = -a.evap_k / r;
                    const float rm = fmaxf(r + 0.5f * h * k1, 1e-9f);
                    r = r + h * (-a.evap_k / rm);
                    if (r <= a.r_min) st = kDead;
                }
            }
            if (st == kDroplet) {
                // leaving the domain kills the droplet
                if (x <= a.lo[0] || x >= a.hi[0] || y <= a.lo[1] || y >= a.hi[1] || z <= a.lo[2] ||
                    z >= a.hi[2])
                    st = kDead;
                else
                    moved = true;
            }
        }

        if (st != st0) a.pstate[p] = st;
        // a droplet that dies in advance_droplets keeps the radius it was
        // given on emission this step (the reference writes pdrop_r before
        // advancing and returns without writing again)
        if (moved)
            a.pdrop_r[p] = r;
        else if (r_new)
            a.pdrop_r[p] = r_emit;
        if (moved) {
            a.px[3 * p] = x;
            a.px[3 * p + 1] = y;
            a.px[3 * p + 2] = z;
            a.pv[3 * p] = vx;
            a.pv[3 * p + 1] = vy;
            a.pv[3 * p + 2] = vz;
        }
        alive = st == kDroplet;
    }
    wave_count(a.counters + 0, emitted);
    wave_count(a.counters + 1, bubble);
    wave_count(a.counters + 2, merged);
    wave_count(a.counters + 3, spray);
    wave_count(a.counters + 4, alive);
}

// |grad u| at cell centres from the MAC faces, with the seven terms the
// reference sums (droplets.py:156-196: dv/dz and dw/dy are not included).
__global__ __launch_bounds__(256) void k_velocity_gradient(const float* __restrict__ u, const float* __restrict__ v,
                                                           const float* __restrict__ w, float* __restrict__ out,
                                                           Dims d, float dx) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nc = (int64_t)d.nx * d.ny * d.nz;
    if (c >= nc) return;
    const int k = (int)(c % d.nz), j = (int)((c / d.nz) % d.ny), i = (int)(c / ((int64_t)d.nz * d.ny));
    auto U = [&](int ii, int jj, int kk) { return u[face_index<0>(d, ii, jj, kk)]; };
    auto V = [&](int ii, int jj, int kk) { return v[face_index<1>(d, ii, jj, kk)]; };
    auto W = [&](int ii, int jj, int kk) { return w[face_index<2>(d, ii, jj, kk)]; };
    const float dudx = (U(i + 1, j, k) - U(i, j, k)) / dx;
    const float dvdy = (V(i, j + 1, k) - V(i, j, k)) / dx;
    const float dwdz = (W(i, j, k + 1) - W(i, j, k)) / dx;
    float acc = dudx * dudx + dvdy * dvdy + dwdz * dwdz;
    const int jm = max(j - 1, 0), jp = min(j + 1, d.ny - 1);
    const int km = max(k - 1, 0), kp = min(k + 1, d.nz - 1);
    const int im = max(i - 1, 0), ip = min(i + 1, d.nx - 1);
    const float dx2 = 2.0f * dx;
    const float dudy = (0.5f * (U(i, jp, k) + U(i + 1, jp, k)) - 0.5f * (U(i, jm, k) + U(i + 1, jm, k))) / dx2;
    const float dudz = (0.5f * (U(i, j, kp) + U(i + 1, j, kp)) - 0.5f * (U(i, j, km) + U(i + 1, j, km))) / dx2;
    const float dvdx = (0.5f * (V(ip, j, k) + V(ip, j + 1, k)) - 0.5f * (V(im, j, k) + V(im, j + 1, k))) / dx2;
    const float dwdx = (0.5f * (W(ip, j, k) + W(ip, j, k + 1)) - 0.5f * (W(im, j, k) + W(im, j, k + 1))) / dx2;
    acc += dudy * dudy + dudz * dudz + dvdx * dvdx + dwdx * dwdx;
    out[c] = sqrtf(acc);
}

// Droplet volume splatted trilinearly into the FP16 spray field
// (droplets.py:308-351): each contribution is rounded to FP16, then added
// with an FP16 atomic.
__global__ __launch_bounds__(256) void k_rasterize_spray(const float* __restrict__ px, const int32_t* __restrict__ pstate,
                                                         const float* __restrict__ pdrop_r, int64_t n,
                                                         __half* __restrict__ dens, Dims d, float dxs) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || pstate[p] != kDroplet) return;
    const float x = px[3 * p], y = px[3 * p + 1], z = px[3 * p + 2];
    const float r = pdrop_r[p];
    const float vol = (4.0f / 3.0f) * 3.14159265f * r * r * r;
    const float gx = x / dxs - 0.5f, gy = y / dxs - 0.5f, gz = z / dxs - 0.5f;
    const int i0 = (int)floorf(gx), j0 = (int)floorf(gy), k0 = (int)floorf(gz);
    const float tx = gx - (float)i0, ty = gy - (float)j0, tz = gz - (float)k0;
    const float cell_vol = dxs * dxs * dxs;
    for (int a = 0; a < 2; ++a) {
        const float wa = a == 0 ? 1.0f - tx : tx;
        const int ii = i0 + a;
        if (ii < 0 || ii >= d.nx) continue;
        for (int b = 0; b < 2; ++b) {
            const float wb = b == 0 ? 1.0f - ty : ty;
            const int jj = j0 + b;
            if (jj < 0 || jj >= d.ny) continue;
            for (int c = 0; c < 2; ++c) {
                const float wc = c == 0 ? 1.0f - tz : tz;
                const int kk = k0 + c;
                if (kk < 0 || kk >= d.nz) continue;
                unsafeAtomicAdd(dens + cell_index(d, ii, jj, kk), __float2half(wa * wb * wc * vol / cell_vol));
            }
        }
    }
}

}  // namespace pf

PF_API int pf_droplets(const pf_droplet_args* in, pf_stream_t stream) {
    PF_REQUIRE(in != nullptr, "pf_droplets: null args");
    PF_REQUIRE(in->n >= 0 && in->n < ((int64_t)1 << 31), "pf_droplets: n out of range");
    PF_REQUIRE(in->nx >= 1 && in->ny >= 1 && in->nz >= 1, "pf_droplets: bad grid");
    PF_REQUIRE(in->counters != nullptr, "pf_droplets: null counters");
    if (in->n == 0) return PF_OK;
    PF_REQUIRE(in->px && in->pv && in->ptype && in->pstate && in->pdrop_r && in->phi,
               "pf_droplets: null particle or phi array");
    PF_REQUIRE(in->vel[0] && in->vel[1] && in->vel[2], "pf_droplets: null velocity");
    PF_REQUIRE(in->max_substeps >= 1, "pf_droplets: max_substeps must be >= 1");
    PF_REQUIRE(in->dx > 0.f, "pf_droplets: dx must be positive");
    pf::DropArgs a;
    a.px = in->px;
    a.pv = in->pv;
    a.ptype = in->ptype;
    a.pstate = in->pstate;
    a.pdrop_r = in->pdrop_r;
    a.n = in->n;
    a.d = pf::Dims{in->nx, in->ny, in->nz};
    a.seed = in->seed;
    a.max_substeps = in->max_substeps;
    a.phi = in->phi;
    a.shear = in->shear;
    for (int q = 0; q < 3; ++q) {
        a.vel[q] = in->vel[q];
        a.g[q] = in->gravity[q];
        a.lo[q] = in->lo[q];
        a.hi[q] = in->hi[q];
    }
    a.dx = in->dx;
    a.dt = in->dt;
    a.delta_phi = in->delta_phi;
    a.r_mu = in->r_mu;
    a.r_sigma = in->r_sigma;
    a.r_min = in->r_min;
    a.r_max = in->r_max;
    a.spray_threshold = in->spray_threshold;
    a.spray_rate = in->spray_rate;
    a.rho_g = in->rho_g;
    a.rho_l = in->rho_l;
    a.cd = in->cd;
    a.evap_k = in->evap_k;
    a.counters = in->counters;
    pf::k_droplets<<<pf::cdiv(in->n, 256), 256, 0, pf::S(stream)>>>(a);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_velocity_gradient(const float* const vel[3], float* out, int nx, int ny, int nz, float dx,
                                pf_stream_t stream) {
    PF_REQUIRE(vel && vel[0] && vel[1] && vel[2] && out, "pf_velocity_gradient: null pointer");
    PF_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "pf_velocity_gradient: bad grid");
    const int64_t nc = (int64_t)nx * ny * nz;
    pf::k_velocity_gradient<<<pf::cdiv(nc, 256), 256, 0, pf::S(stream)>>>(vel[0], vel[1], vel[2], out,
                                                                           pf::Dims{nx, ny, nz}, dx);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_rasterize_spray(const float* px, const int32_t* pstate, const float* pdrop_r, int64_t n,
                              uint16_t* density, int sx, int sy, int sz, float dx_spray, pf_stream_t stream) {
    PF_REQUIRE(n >= 0 && n < ((int64_t)1 << 31), "pf_rasterize_spray: n out of range");
    PF_REQUIRE(sx >= 1 && sy >= 1 && sz >= 1 && dx_spray > 0.f, "pf_rasterize_spray: bad grid");
    PF_REQUIRE(density != nullptr, "pf_rasterize_spray: null density");
    PF_REQUIRE((reinterpret_cast<uintptr_t>(density) & 3) == 0, "pf_rasterize_spray: density must be 4-byte aligned");
    if (n == 0) return PF_OK;
    PF_REQUIRE(px && pstate && pdrop_r, "pf_rasterize_spray: null particle array");
    pf::k_rasterize_spray<<<pf::cdiv(n, 256), 256, 0, pf::S(stream)>>>(
        px, pstate, pdrop_r, n, reinterpret_cast<__half*>(density), pf::Dims{sx, sy, sz}, dx_spray);
    PF_CHECK_LAUNCH();
    return PF_OK;
}
