// Hand-written adjoint (reverse mode) of one PF-FLIP step.
//
// The forward step (pf_p2g -> pf_phase_beta_forces -> pf_mgpcg_solve ->
// pf_project -> pf_extrapolate -> pf_g2p_advect) maps the particle state
// (x, v) to (x', v').  Given the adjoints (x', v')-bar, these kernels return
// (x, v)-bar, with dt held fixed.  Each kernel is the transpose of one
// forward stage, evaluated at the forward state that the caller rebuilds
// from a per-step checkpoint:
//   G2P + RK3 + reflect  -> per-particle recompute, then a P2G-shaped scatter
//                           of the grid adjoints (pf_adj_g2p_advect)
//   extrapolation        -> transposed Jacobi sweeps (pf_adj_extrapolate)
//   projection           -> cell gather of p-bar, face beta-bar (pf_adj_project)
//   pressure solve       -> the same symmetric MG-PCG on p-bar (caller), then
//                           beta-bar -= (lam_i - lam_j)(p_i - p_j) (pf_adj_solve_beta)
//   RHS / gravity / walls / vel_star / eq. (7) / beta -> per face (pf_adj_phase);
//                           the eq. (7) clamps use their one-sided subgradients
//   P2G                  -> a G2P-shaped gather per particle (pf_adj_p2g)
// The cell mass only feeds the liquid tags (a threshold), so it carries no
// gradient; the tags, wall/solid masks and extrapolation masks are held at
// their forward values.
#include "pf_common.hpp"

namespace pf {

// ------------------------------------------------------------------ trilinear taps with gradient
// Trilinear interpolation of the face grid of axis A at grid coordinates
// g = x/dx, indices clamped per axis (kernels.py:52-79), together with the
// derivative of the interpolated value with respect to x.
template <int A>
struct TapsG {
    int idx[8];
    float w[8];
    float dw[3][8];  // d w / d x (per unit length)
};

template <int A>
__device__ inline void make_taps_g(float gx, float gy, float gz, float inv_dx, const Dims& d, TapsG<A>& t) {
    const float ox = A == 0 ? 0.f : 0.5f, oy = A == 1 ? 0.f : 0.5f, oz = A == 2 ? 0.f : 0.5f;
    const int n0 = d.nx + (A == 0), n1 = d.ny + (A == 1), n2 = d.nz + (A == 2);
    const float hx = gx - ox, hy = gy - oy, hz = gz - oz;
    const int i0 = (int)floorf(hx), j0 = (int)floorf(hy), k0 = (int)floorf(hz);
    const float fx = hx - (float)i0, fy = hy - (float)j0, fz = hz - (float)k0;
#pragma unroll
    for (int di = 0; di < 2; ++di) {
        const float wi = di == 0 ? 1.0f - fx : fx, si = di == 0 ? -inv_dx : inv_dx;
        const int ii = min(max(i0 + di, 0), n0 - 1);
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const float wj = dj == 0 ? 1.0f - fy : fy, sj = dj == 0 ? -inv_dx : inv_dx;
            const int jj = min(max(j0 + dj, 0), n1 - 1);
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                const float wk = dk == 0 ? 1.0f - fz : fz, sk = dk == 0 ? -inv_dx : inv_dx;
                const int kk = min(max(k0 + dk, 0), n2 - 1);
                const int q = di * 4 + dj * 2 + dk;
                t.w[q] = wi * wj * wk;
                t.dw[0][q] = si * wj * wk;
                t.dw[1][q] = wi * sj * wk;
                t.dw[2][q] = wi * wj * sk;
                t.idx[q] = (ii * n1 + jj) * n2 + kk;
            }
        }
    }
}

template <int A>
__device__ inline float interp(const TapsG<A>& t, const float* __restrict__ f) {
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) acc += t.w[q] * f[t.idx[q]];
    return acc;
}

// gbar += dI/dx^T s  (s: adjoint of the interpolated value)
template <int A>
__device__ inline void interp_grad(const TapsG<A>& t, const float* __restrict__ f, float s, float (&gbar)[3]) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float v = f[t.idx[q]];
        a0 += t.dw[0][q] * v;
        a1 += t.dw[1][q] * v;
        a2 += t.dw[2][q] * v;
    }
    gbar[0] += a0 * s;
    gbar[1] += a1 * s;
    gbar[2] += a2 * s;
}

// field-bar += I^T s (scatter)
template <int A>
__device__ inline void interp_scatter(const TapsG<A>& t, float* f_bar, float s) {
    if (s == 0.0f) return;
#pragma unroll
    for (int q = 0; q < 8; ++q) unsafeAtomicAdd(&f_bar[t.idx[q]], t.w[q] * s);
}

// One velocity sample U(y) = (I_0 f_0, I_1 f_1, I_2 f_2)(y), and its adjoint:
// ybar += J^T kbar, fbar += I^T kbar.
struct VelSample {
    TapsG<0> t0;
    TapsG<1> t1;
    TapsG<2> t2;
};

__device__ inline void vel_sample(VelSample& s, const float* const* f, float x, float y, float z, float dx,
                                  const Dims& d, float (&k)[3]) {
    const float gx = x / dx, gy = y / dx, gz = z / dx, inv = 1.0f / dx;
    make_taps_g<0>(gx, gy, gz, inv, d, s.t0);
    make_taps_g<1>(gx, gy, gz, inv, d, s.t1);
    make_taps_g<2>(gx, gy, gz, inv, d, s.t2);
    k[0] = interp(s.t0, f[0]);
    k[1] = interp(s.t1, f[1]);
    k[2] = interp(s.t2, f[2]);
}

__device__ inline void vel_sample_adj(const VelSample& s, const float* const* f, float* const* fbar,
                                      const float (&kbar)[3], float (&ybar)[3]) {
    interp_grad(s.t0, f[0], kbar[0], ybar);
    interp_grad(s.t1, f[1], kbar[1], ybar);
    interp_grad(s.t2, f[2], kbar[2], ybar);
    interp_scatter(s.t0, fbar[0], kbar[0]);
    interp_scatter(s.t1, fbar[1], kbar[1]);
    interp_scatter(s.t2, fbar[2], kbar[2]);
}

// ------------------------------------------------------------------ G2P + RK3 + reflect
struct AdjG2PArgs {
    const float* px;  // forward input x (pre-step)
    const float* pv;  // forward input v
    const int* ptype;
    const int* pstate;
    int64_t n;
    Dims d;
    int flags;
    const float* vel[3];
    const float* dvel[3];
    const float* vel_liq[3];
    const int* solid;
    float dx, dt, alpha_l, alpha_a;
    float lo[3], hi[3];
    const float* gx_out;  // adjoint of x'
    const float* gv_out;  // adjoint of v'
    float* gx_in;         // adjoint of x (written)
    float* gv_in;         // adjoint of v (written)
    float* g_vel[3];      // accumulated (atomics)
    float* g_dvel[3];
    float* g_vel_liq[3];
};

__global__ __launch_bounds__(256) void k_adj_g2p_advect(AdjG2PArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const int st = a.pstate[p];
    const bool liquid = a.ptype[p] == 1, active = st == 0, dead = st == 2;
    float gx[3] = {0.f, 0.f, 0.f}, gv[3] = {0.f, 0.f, 0.f};
    if (dead) {
        // dead particles are not touched by the forward pass
        for (int q = 0; q < 3; ++q) {
            a.gx_in[3 * p + q] = a.gx_out[3 * p + q];
            a.gv_in[3 * p + q] = a.gv_out[3 * p + q];
        }
        return;
    }
    const Dims d = a.d;
    const float dx = a.dx, dt = a.dt;
    const float x0[3] = {a.px[3 * p], a.px[3 * p + 1], a.px[3 * p + 2]};
    const float v0[3] = {a.pv[3 * p], a.pv[3 * p + 1], a.pv[3 * p + 2]};

    // ---- forward recompute
    VelSample sx;  // samples at x: vel and dvel share the taps
    float pic[3] = {0.f, 0.f, 0.f}, dux[3] = {0.f, 0.f, 0.f};
    const bool do_g2p = active && (a.flags & PF_GA_G2P);
    const bool do_adv = active && (a.flags & PF_GA_ADVECT);
    const float al = liquid ? a.alpha_l : a.alpha_a;
    float v1[3] = {v0[0], v0[1], v0[2]};
    if (do_g2p || do_adv) vel_sample(sx, a.vel, x0[0], x0[1], x0[2], dx, d, pic);
    if (do_g2p) {
        dux[0] = interp(sx.t0, a.dvel[0]);
        dux[1] = interp(sx.t1, a.dvel[1]);
        dux[2] = interp(sx.t2, a.dvel[2]);
        for (int q = 0; q < 3; ++q) v1[q] = al * (v0[q] + dux[q]) + (1.0f - al) * pic[q];
    }
    const float* const* f = liquid ? a.vel_liq : a.vel;
    float* const* fbar = liquid ? a.g_vel_liq : a.g_vel;
    VelSample s1, s2, s3;
    float k1[3], k2[3], k3[3], xc[3] = {x0[0], x0[1], x0[2]}, y2[3], y3[3];
    bool clamped[3] = {false, false, false};
    const float h1 = 0.5f * dt, h2 = 0.75f * dt;
    const float c1 = 2.0f / 9.0f, c2 = 3.0f / 9.0f, c3 = 4.0f / 9.0f;
    if (do_adv) {
        if (liquid)
            vel_sample(s1, f, x0[0], x0[1], x0[2], dx, d, k1);
        else
            for (int q = 0; q < 3; ++q) k1[q] = pic[q];
        for (int q = 0; q < 3; ++q) y2[q] = x0[q] + h1 * k1[q];
        vel_sample(s2, f, y2[0], y2[1], y2[2], dx, d, k2);
        for (int q = 0; q < 3; ++q) y3[q] = x0[q] + h2 * k2[q];
        vel_sample(s3, f, y3[0], y3[1], y3[2], dx, d, k3);
        for (int q = 0; q < 3; ++q) {
            const float xn = x0[q] + dt * (c1 * k1[q] + c2 * k2[q] + c3 * k3[q]);
            xc[q] = fminf(fmaxf(xn, a.lo[q]), a.hi[q]);
            clamped[q] = xc[q] != xn;
        }
    }

    // ---- backward
    float xb[3], vb[3];
    for (int q = 0; q < 3; ++q) {
        xb[q] = a.gx_out[3 * p + q];
        vb[q] = a.gv_out[3 * p + q];
    }
    if ((a.flags & PF_GA_SOLIDS) && a.solid) {
        // push_out_of_solids (kernels.py:630-679) after the reflection: the
        // position moves by a constant, and an inward velocity component is
        // zeroed (its adjoint with it)
        float xr[3], vr[3];
        for (int q = 0; q < 3; ++q) {
            xr[q] = xc[q];
            vr[q] = v1[q];
            if (a.flags & PF_GA_REFLECT) {
                if (xr[q] <= a.lo[q]) {
                    xr[q] = a.lo[q];
                    vr[q] = fmaxf(vr[q], 0.f);
                } else if (xr[q] >= a.hi[q]) {
                    xr[q] = a.hi[q];
                    vr[q] = fminf(vr[q], 0.f);
                }
            }
        }
        const int i = min(max((int)(xr[0] / dx), 0), d.nx - 1);
        const int j = min(max((int)(xr[1] / dx), 0), d.ny - 1);
        const int k = min(max((int)(xr[2] / dx), 0), d.nz - 1);
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
                        if ((float)t < best) {
                            best = (float)t;
                            bax = ax;
                            bsg = sg;
                        }
                        break;
                    }
                }
            }
            if (best <= 1.0e8f && (float)bsg * vr[bax] < 0.f) vb[bax] = 0.0f;
        }
    }
    if (a.flags & PF_GA_REFLECT) {
        // kernels.py:567-595: a particle on a wall plane keeps that coordinate
        // fixed and loses the inward velocity component
        for (int q = 0; q < 3; ++q) {
            if (xc[q] <= a.lo[q]) {
                xb[q] = 0.0f;
                if (!(v1[q] > 0.0f)) vb[q] = 0.0f;
            } else if (xc[q] >= a.hi[q]) {
                xb[q] = 0.0f;
                if (!(v1[q] < 0.0f)) vb[q] = 0.0f;
            }
        }
    }
    for (int q = 0; q < 3; ++q) gx[q] = xb[q];
    if (do_adv) {
        float kb1[3], kb2[3], kb3[3];
        for (int q = 0; q < 3; ++q) {
            if (clamped[q]) xb[q] = 0.0f;  // clamp into [lo, hi]
            gx[q] = xb[q];
            kb1[q] = dt * c1 * xb[q];
            kb2[q] = dt * c2 * xb[q];
            kb3[q] = dt * c3 * xb[q];
        }
        float yb[3] = {0.f, 0.f, 0.f};
        vel_sample_adj(s3, f, fbar, kb3, yb);  // k3 = U(y3), y3 = x + h2 k2
        for (int q = 0; q < 3; ++q) {
            gx[q] += yb[q];
            kb2[q] += h2 * yb[q];
            yb[q] = 0.f;
        }
        vel_sample_adj(s2, f, fbar, kb2, yb);  // k2 = U(y2), y2 = x + h1 k1
        for (int q = 0; q < 3; ++q) {
            gx[q] += yb[q];
            kb1[q] += h1 * yb[q];
            yb[q] = 0.f;
        }
        if (liquid) {
            vel_sample_adj(s1, f, fbar, kb1, yb);  // k1 = U(x)
        } else {
            vel_sample_adj(sx, a.vel, a.g_vel, kb1, yb);  // air: k1 is the PIC sample I(vel)(x)
        }
        for (int q = 0; q < 3; ++q) gx[q] += yb[q];
    }
    if (do_g2p) {
        // v1 = al (v0 + I(dvel)) + (1 - al) I(vel)
        float yb[3] = {0.f, 0.f, 0.f};
        const float sd[3] = {al * vb[0], al * vb[1], al * vb[2]};
        const float sp[3] = {(1.0f - al) * vb[0], (1.0f - al) * vb[1], (1.0f - al) * vb[2]};
        vel_sample_adj(sx, a.dvel, a.g_dvel, sd, yb);
        vel_sample_adj(sx, a.vel, a.g_vel, sp, yb);
        for (int q = 0; q < 3; ++q) {
            gx[q] += yb[q];
            gv[q] = al * vb[q];
        }
    } else {
        for (int q = 0; q < 3; ++q) gv[q] = vb[q];
    }
    for (int q = 0; q < 3; ++q) {
        a.gx_in[3 * p + q] = gx[q];
        a.gv_in[3 * p + q] = gv[q];
    }
}

// ------------------------------------------------------------------ extrapolation
// Forward levels: 0 = seeded (touches a liquid cell), s = set by sweep s,
// 255 = never set (value 0).  Same rules as k_extrap_seed / k_extrap_sweep.
struct AdjExtrapK {
    Dims d;
    const int* is_liq;
    uint8_t* lvl[3];
    float* g_liq[3];  // adjoint of vel_liq, consumed in place
    float* g_vel[3];  // adjoint of vel, accumulated
    int s;            // sweep being processed
};

__device__ inline bool face_of(const Dims& d, int64_t t, int& A, int& i, int& j, int& k, int& f) {
    const int64_t f0 = face_count(d, 0), f1 = face_count(d, 1), f2 = face_count(d, 2);
    if (t < f0) {
        A = 0;
    } else if (t < f0 + f1) {
        A = 1;
        t -= f0;
    } else if (t < f0 + f1 + f2) {
        A = 2;
        t -= f0 + f1;
    } else {
        return false;
    }
    const int sy = d.ny + (A == 1), sz = d.nz + (A == 2);
    f = (int)t;
    k = (int)(t % sz);
    j = (int)((t / sz) % sy);
    i = (int)(t / ((int64_t)sz * sy));
    return true;
}

__global__ __launch_bounds__(256) void k_adj_extrap_seed(AdjExtrapK a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int A, i, j, k, f;
    if (!face_of(a.d, t, A, i, j, k, f)) return;
    const Dims d = a.d;
    int liq = 0;
    if (i < d.nx && j < d.ny && k < d.nz) liq = max(liq, a.is_liq[cell_index(d, i, j, k)]);
    const int im = i - (A == 0), jm = j - (A == 1), km = k - (A == 2);
    if (im >= 0 && jm >= 0 && km >= 0) liq = max(liq, a.is_liq[cell_index(d, im, jm, km)]);
    a.lvl[A][f] = liq == 1 ? 0 : 255;
}

// number of neighbours of face (i,j,k) with level < s
__device__ inline int known_count(const uint8_t* lv, int sx, int sy, int sz, int i, int j, int k, int s) {
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int ax = q / 2, sg = (q % 2 == 0) ? -1 : 1;
        const int ii = i + (ax == 0 ? sg : 0), jj = j + (ax == 1 ? sg : 0), kk = k + (ax == 2 ? sg : 0);
        if (ii < 0 || jj < 0 || kk < 0 || ii >= sx || jj >= sy || kk >= sz) continue;
        if (lv[(ii * sy + jj) * sz + kk] < s) ++cnt;
    }
    return cnt;
}

// forward sweep s: unknown faces with a known neighbour get level s
__global__ __launch_bounds__(256) void k_adj_extrap_level(AdjExtrapK a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int A, i, j, k, f;
    if (!face_of(a.d, t, A, i, j, k, f)) return;
    uint8_t* lv = a.lvl[A];
    if (lv[f] != 255) return;
    const Dims d = a.d;
    const int sx = d.nx + (A == 0), sy = d.ny + (A == 1), sz = d.nz + (A == 2);
    // faces set in this sweep read only levels < s, so writing s is race-free
    if (known_count(lv, sx, sy, sz, i, j, k, a.s) > 0) lv[f] = (uint8_t)a.s;
}

// backward sweep s: every face with level < s collects g/cnt from its
// neighbours of level s (their value was the mean of their cnt neighbours
// of level < s)
__global__ __launch_bounds__(256) void k_adj_extrap_sweep(AdjExtrapK a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int A, i, j, k, f;
    if (!face_of(a.d, t, A, i, j, k, f)) return;
    const uint8_t* lv = a.lvl[A];
    if (lv[f] >= a.s) return;
    const Dims d = a.d;
    const int sx = d.nx + (A == 0), sy = d.ny + (A == 1), sz = d.nz + (A == 2);
    float acc = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int ax = q / 2, sg = (q % 2 == 0) ? -1 : 1;
        const int ii = i + (ax == 0 ? sg : 0), jj = j + (ax == 1 ? sg : 0), kk = k + (ax == 2 ? sg : 0);
        if (ii < 0 || jj < 0 || kk < 0 || ii >= sx || jj >= sy || kk >= sz) continue;
        const int g = (ii * sy + jj) * sz + kk;
        if (lv[g] == a.s) acc += a.g_liq[A][g] / (float)known_count(lv, sx, sy, sz, ii, jj, kk, a.s);
    }
    if (acc != 0.0f) a.g_liq[A][f] += acc;
}

// seeded faces copy vel: vel-bar += vel_liq-bar there
__global__ __launch_bounds__(256) void k_adj_extrap_final(AdjExtrapK a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int A, i, j, k, f;
    if (!face_of(a.d, t, A, i, j, k, f)) return;
    if (a.lvl[A][f] == 0) a.g_vel[A][f] += a.g_liq[A][f];
}

// ------------------------------------------------------------------ projection
// vel3 = vel2 - dt beta (p_hi - p_lo) / dx on interior faces with beta != 0;
// dvel = vel3 - vel_star.  In: g_vel (vel3-bar from G2P / advection /
// extrapolation), g_dvel.  Out: g_vel <- vel2-bar = vel3-bar + dvel-bar,
// g_dvel <- vel_star-bar = -dvel-bar, g_beta (overwritten), g_p (overwritten).
struct AdjProjK {
    Dims d;
    const float* p;
    const float* beta[3];
    float* g_vel[3];
    float* g_dvel[3];
    float* g_beta[3];
    float* g_p;
    float dx, dt;
};

template <int A>
__device__ inline float adj_proj_face(const AdjProjK& a, int i, int j, int k) {
    const Dims d = a.d;
    const int f = face_index<A>(d, i, j, k);
    const float ub = a.g_vel[A][f] + a.g_dvel[A][f];
    a.g_vel[A][f] = ub;
    a.g_dvel[A][f] = -a.g_dvel[A][f];
    const int ia = A == 0 ? i : (A == 1 ? j : k);
    const int na = A == 0 ? d.nx : (A == 1 ? d.ny : d.nz);
    const float b = a.beta[A][f];
    float gb = 0.0f;
    if (b != 0.0f && ia > 0 && ia < na) {
        const float ph = a.p[cell_index(d, i, j, k)];
        const float pl = a.p[cell_index(d, i - (A == 0), j - (A == 1), k - (A == 2))];
        gb = -(a.dt / a.dx) * (ph - pl) * ub;
    }
    a.g_beta[A][f] = gb;
    return ub;
}

// p-bar of cell c: -(dt/dx) beta ub over its low faces, + over its high faces
template <int A>
__device__ inline float adj_proj_cell(const AdjProjK& a, int i, int j, int k, bool low) {
    const Dims d = a.d;
    const int ii = i + (!low && A == 0), jj = j + (!low && A == 1), kk = k + (!low && A == 2);
    const int ia = A == 0 ? ii : (A == 1 ? jj : kk);
    const int na = A == 0 ? d.nx : (A == 1 ? d.ny : d.nz);
    if (ia <= 0 || ia >= na) return 0.0f;
    const int f = face_index<A>(d, ii, jj, kk);
    const float b = a.beta[A][f];
    if (b == 0.0f) return 0.0f;
    const float ub = a.g_vel[A][f];  // already vel2-bar == vel3-bar
    const float s = (a.dt / a.dx) * b * ub;
    return low ? -s : s;
}

__global__ __launch_bounds__(256) void k_adj_project_faces(AdjProjK a) {
    const Dims d = a.d;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ez = d.nz + 1, ey = d.ny + 1;
    if (t >= (int64_t)(d.nx + 1) * ey * ez) return;
    const int k = (int)(t % ez), j = (int)((t / ez) % ey), i = (int)(t / ((int64_t)ez * ey));
    const bool ci = i < d.nx, cj = j < d.ny, ck = k < d.nz;
    if (cj && ck) adj_proj_face<0>(a, i, j, k);
    if (ci && ck) adj_proj_face<1>(a, i, j, k);
    if (ci && cj) adj_proj_face<2>(a, i, j, k);
}

__global__ __launch_bounds__(256) void k_adj_project_cells(AdjProjK a) {
    const Dims d = a.d;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int64_t)d.nx * d.ny * d.nz) return;
    const int k = (int)(c % d.nz), j = (int)((c / d.nz) % d.ny), i = (int)(c / ((int64_t)d.nz * d.ny));
    float s = adj_proj_cell<0>(a, i, j, k, true) + adj_proj_cell<0>(a, i, j, k, false);
    s += adj_proj_cell<1>(a, i, j, k, true) + adj_proj_cell<1>(a, i, j, k, false);
    s += adj_proj_cell<2>(a, i, j, k, true) + adj_proj_cell<2>(a, i, j, k, false);
    a.g_p[c] = s;
}

// ------------------------------------------------------------------ solve: beta-bar
// p solves A(beta) p = b with (A p)_i = sum_f beta_f (p_i - p_nbr); with
// lam = A^+ p-bar:  beta_f-bar -= (lam_i - lam_j)(p_i - p_j) (interior faces)
__global__ __launch_bounds__(256) void k_adj_solve_beta(Dims d, const float* __restrict__ p,
                                                        const float* __restrict__ lam, float* gx, float* gy,
                                                        float* gz) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ez = d.nz + 1, ey = d.ny + 1;
    if (t >= (int64_t)(d.nx + 1) * ey * ez) return;
    const int k = (int)(t % ez), j = (int)((t / ez) % ey), i = (int)(t / ((int64_t)ez * ey));
    const bool ci = i < d.nx, cj = j < d.ny, ck = k < d.nz;
    if (ci && cj && ck) {
        const int c = cell_index(d, i, j, k);
        if (i > 0) {
            const int o = cell_index(d, i - 1, j, k);
            gx[face_index<0>(d, i, j, k)] -= (lam[c] - lam[o]) * (p[c] - p[o]);
        }
        if (j > 0) {
            const int o = cell_index(d, i, j - 1, k);
            gy[face_index<1>(d, i, j, k)] -= (lam[c] - lam[o]) * (p[c] - p[o]);
        }
        if (k > 0) {
            const int o = cell_index(d, i, j, k - 1);
            gz[face_index<2>(d, i, j, k)] -= (lam[c] - lam[o]) * (p[c] - p[o]);
        }
    }
}

// ------------------------------------------------------------------ phase / forces / RHS
struct AdjPhaseK {
    Dims d;
    const float* mass[3];
    const float* mom[3];
    const float* beta[3];
    const int* is_liq;
    const int* solid;
    const float* g_vel2[3];     // vel2-bar (after projection adjoint)
    const float* g_vel_star[3];
    const float* g_beta[3];
    const float* g_rhs;         // lam
    float* g_mass[3];           // written
    float* g_mom[3];            // written
    float mass_eps, rho_min, alpha_phi, rho_0l, rho_l, rho_g, dt, dx;
};

template <int A>
__device__ inline void adj_phase_face(const AdjPhaseK& a, int i, int j, int k) {
    const Dims d = a.d;
    const int f = face_index<A>(d, i, j, k);
    const int ia = A == 0 ? i : (A == 1 ? j : k);
    const int na = A == 0 ? d.nx : (A == 1 ? d.ny : d.nz);
    const bool lo = ia == 0, hi = ia == na;
    const int clo = lo ? -1 : cell_index(d, i - (A == 0), j - (A == 1), k - (A == 2));
    const int chi = hi ? -1 : cell_index(d, i, j, k);
    const bool blocked = lo || hi || (a.solid && (a.solid[clo] == 1 || a.solid[chi] == 1));
    // rhs(c) = -(dx/dt) sum (u_hi - u_lo): this face is u_lo of chi, u_hi of clo
    float ub = a.g_vel2[A][f];
    const float r = a.dx / a.dt;
    if (!hi) ub += r * a.g_rhs[chi];
    if (!lo) ub -= r * a.g_rhs[clo];
    float u1 = blocked ? 0.0f : ub;  // gravity, then re-zero
    u1 += a.g_vel_star[A][f];        // vel_star = vel after the first zeroing
    const float u0 = blocked ? 0.0f : u1;
    const float m = a.mass[A][f];
    float gm = 0.0f, gu = 0.0f;
    if (m > a.mass_eps) {
        gu = u0 / m;
        gm = -u0 * (a.mom[A][f] / m) / m;
    }
    if (!blocked) {
        // beta = 1 / ((1 - phi) rho_g + phi rho_l): d beta / d phi = -beta^2 (rho_l - rho_g)
        const float b = a.beta[A][f];
        const float gphi = a.g_beta[A][f] * (-b * b) * (a.rho_l - a.rho_g);
        const int ta = a.is_liq[clo], tb = a.is_liq[chi];
        if (ta != tb && m >= a.rho_min) {
            // eq. (7) on the face mass; zero slope where it is clamped at 1
            const float s = fmaxf(m - a.rho_min, 0.0f) / (a.alpha_phi * a.rho_0l);
            const float phi = sqrtf(s);
            if (phi > 0.0f && phi < 1.0f) gm += gphi * 0.5f / (phi * a.alpha_phi * a.rho_0l);
        }
    }
    a.g_mass[A][f] = gm;
    a.g_mom[A][f] = gu;
}

__global__ __launch_bounds__(256) void k_adj_phase(AdjPhaseK a) {
    const Dims d = a.d;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int ez = d.nz + 1, ey = d.ny + 1;
    if (t >= (int64_t)(d.nx + 1) * ey * ez) return;
    const int k = (int)(t % ez), j = (int)((t / ez) % ey), i = (int)(t / ((int64_t)ez * ey));
    const bool ci = i < d.nx, cj = j < d.ny, ck = k < d.nz;
    if (cj && ck) adj_phase_face<0>(a, i, j, k);
    if (ci && ck) adj_phase_face<1>(a, i, j, k);
    if (ci && cj) adj_phase_face<2>(a, i, j, k);
}

// ------------------------------------------------------------------ P2G
// mass_f = sum_p w_fp m_p, mom_f = sum_p w_fp m_p v_p,a (eq. 6, 2 candidates
// per axis): v-bar_a += sum_f w m mom_f-bar, x-bar += sum_f m (mass_f-bar +
// v_a mom_f-bar) grad w.  A gather over the particle's candidate faces.
struct AdjP2GK {
    const float* px;
    const float* pv;
    const float* pmass;
    const int* pstate;
    int64_t n;
    Dims d;
    float dx, inv_r2;
    const float* g_mass[3];
    const float* g_mom[3];
    float* gx;  // accumulated
    float* gv;  // accumulated
};

template <int A>
__device__ inline void adj_p2g_axis(const AdjP2GK& a, const float (&x)[3], float v, float m, float (&gx)[3],
                                    float& gv) {
    const Dims d = a.d;
    const float dx = a.dx;
    const int n0 = d.nx + (A == 0), n1 = d.ny + (A == 1), n2 = d.nz + (A == 2);
    int base[3];
    float node0[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const float g = x[q] / dx;
        if (q == A) {
            base[q] = (int)floorf(g);  // normal: faces floor(g), floor(g)+1
            node0[q] = (float)base[q] * dx;
        } else {
            base[q] = (int)floorf(g - 0.5f);  // tangential: centres floor(g-1/2) (+1)
            node0[q] = ((float)base[q] + 0.5f) * dx;
        }
    }
    const float* gm = a.g_mass[A];
    const float* gu = a.g_mom[A];
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                const int ii = base[0] + di, jj = base[1] + dj, kk = base[2] + dk;
                if (ii < 0 || jj < 0 || kk < 0 || ii >= n0 || jj >= n1 || kk >= n2) continue;
                const float r0 = x[0] - (node0[0] + (float)di * dx);
                const float r1 = x[1] - (node0[1] + (float)dj * dx);
                const float r2 = x[2] - (node0[2] + (float)dk * dx);
                const float tt = 1.0f - (r0 * r0 + r1 * r1 + r2 * r2) * a.inv_r2;
                if (!(tt > 0.0f)) continue;
                const float w = tt * tt * tt;
                const int fi = (ii * n1 + jj) * n2 + kk;
                const float bm = gm[fi], bu = gu[fi];
                gv += w * m * bu;
                // d w / d x = 3 t^2 * (-2 r / r_p^2)
                const float s = m * (bm + v * bu) * 3.0f * tt * tt * (-2.0f * a.inv_r2);
                gx[0] += s * r0;
                gx[1] += s * r1;
                gx[2] += s * r2;
            }
}

__global__ __launch_bounds__(256) void k_adj_p2g(AdjP2GK a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    if (a.pstate[p] != 0) return;
    const float x[3] = {a.px[3 * p], a.px[3 * p + 1], a.px[3 * p + 2]};
    const float m = a.pmass[p];
    float gx[3] = {0.f, 0.f, 0.f}, gv[3] = {0.f, 0.f, 0.f};
    adj_p2g_axis<0>(a, x, a.pv[3 * p], m, gx, gv[0]);
    adj_p2g_axis<1>(a, x, a.pv[3 * p + 1], m, gx, gv[1]);
    adj_p2g_axis<2>(a, x, a.pv[3 * p + 2], m, gx, gv[2]);
    for (int q = 0; q < 3; ++q) {
        a.gx[3 * p + q] += gx[q];
        a.gv[3 * p + q] += gv[q];
    }
}

}  // namespace pf

using namespace pf;

PF_API int pf_adj_g2p_advect(const pf_g2p_args* fwd, const float* gx_out, const float* gv_out, float* gx_in,
                             float* gv_in, float* const g_vel[3], float* const g_dvel[3],
                             float* const g_vel_liq[3], pf_stream_t stream) {
    PF_REQUIRE(fwd, "pf_adj_g2p_advect: null args");
    const pf_g2p_args& in = *fwd;
    PF_REQUIRE(in.px && in.pv && in.ptype && in.pstate, "pf_adj_g2p_advect: need px, pv, ptype, pstate");
    PF_REQUIRE(!in.pflag, "pf_adj_g2p_advect: packed phase flags are not supported");
    PF_REQUIRE(!(in.flags & PF_GA_SOLIDS) || in.solid, "pf_adj_g2p_advect: solids flag without mask");
    PF_REQUIRE(gx_out && gv_out && gx_in && gv_in && g_vel && g_dvel && g_vel_liq, "pf_adj_g2p_advect: null adjoint");
    PF_REQUIRE(in.nx > 0 && in.ny > 0 && in.nz > 0 && in.n >= 0 && in.dx > 0.0f, "pf_adj_g2p_advect: bad sizes");
    AdjG2PArgs a;
    a.px = in.px;
    a.pv = in.pv;
    a.ptype = in.ptype;
    a.pstate = in.pstate;
    a.n = in.n;
    a.d = Dims{in.nx, in.ny, in.nz};
    a.flags = in.flags;
    for (int q = 0; q < 3; ++q) {
        PF_REQUIRE(in.vel[q] && g_vel[q], "pf_adj_g2p_advect: need vel");
        if (in.flags & PF_GA_G2P) PF_REQUIRE(in.dvel[q] && g_dvel[q], "pf_adj_g2p_advect: need dvel");
        if (in.flags & PF_GA_ADVECT) PF_REQUIRE(in.vel_liq[q] && g_vel_liq[q], "pf_adj_g2p_advect: need vel_liq");
        a.vel[q] = in.vel[q];
        a.dvel[q] = in.dvel[q];
        a.vel_liq[q] = in.vel_liq[q];
        a.g_vel[q] = g_vel[q];
        a.g_dvel[q] = g_dvel[q];
        a.g_vel_liq[q] = g_vel_liq[q];
        a.lo[q] = in.lo[q];
        a.hi[q] = in.hi[q];
    }
    a.solid = in.solid;
    a.dx = in.dx;
    a.dt = in.dt;
    a.alpha_l = in.alpha_liquid;
    a.alpha_a = in.alpha_air;
    a.gx_out = gx_out;
    a.gv_out = gv_out;
    a.gx_in = gx_in;
    a.gv_in = gv_in;
    if (in.n == 0) return PF_OK;
    k_adj_g2p_advect<<<cdiv(in.n, 256), 256, 0, S(stream)>>>(a);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_adj_extrapolate(const int32_t* is_liquid, float* const g_vel_liq[3], float* const g_vel[3],
                              uint8_t* const level[3], int nx, int ny, int nz, int sweeps, pf_stream_t stream) {
    PF_REQUIRE(nx > 0 && ny > 0 && nz > 0, "pf_adj_extrapolate: bad grid");
    PF_REQUIRE(sweeps >= 0 && sweeps < 255, "pf_adj_extrapolate: bad sweep count");
    PF_REQUIRE(is_liquid && g_vel_liq && g_vel && level, "pf_adj_extrapolate: null argument");
    AdjExtrapK a;
    a.d = Dims{nx, ny, nz};
    a.is_liq = is_liquid;
    for (int q = 0; q < 3; ++q) {
        PF_REQUIRE(g_vel_liq[q] && g_vel[q] && level[q], "pf_adj_extrapolate: null argument");
        a.lvl[q] = level[q];
        a.g_liq[q] = g_vel_liq[q];
        a.g_vel[q] = g_vel[q];
    }
    const int64_t n = face_count(a.d, 0) + face_count(a.d, 1) + face_count(a.d, 2);
    const unsigned grid = cdiv(n, 256);
    hipStream_t s = S(stream);
    a.s = 0;
    k_adj_extrap_seed<<<grid, 256, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    for (int sw = 1; sw <= sweeps; ++sw) {
        a.s = sw;
        k_adj_extrap_level<<<grid, 256, 0, s>>>(a);
        PF_CHECK_LAUNCH();
    }
    for (int sw = sweeps; sw >= 1; --sw) {
        a.s = sw;
        k_adj_extrap_sweep<<<grid, 256, 0, s>>>(a);
        PF_CHECK_LAUNCH();
    }
    k_adj_extrap_final<<<grid, 256, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_adj_project(const float* p, float* const beta[3], float* const g_vel[3], float* const g_dvel[3],
                          float* const g_beta[3], float* g_p, int nx, int ny, int nz, float dx, float dt,
                          pf_stream_t stream) {
    PF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && dx > 0.0f, "pf_adj_project: bad grid");
    PF_REQUIRE(p && beta && g_vel && g_dvel && g_beta && g_p, "pf_adj_project: null argument");
    AdjProjK a;
    a.d = Dims{nx, ny, nz};
    a.p = p;
    for (int q = 0; q < 3; ++q) {
        PF_REQUIRE(beta[q] && g_vel[q] && g_dvel[q] && g_beta[q], "pf_adj_project: null argument");
        a.beta[q] = beta[q];
        a.g_vel[q] = g_vel[q];
        a.g_dvel[q] = g_dvel[q];
        a.g_beta[q] = g_beta[q];
    }
    a.g_p = g_p;
    a.dx = dx;
    a.dt = dt;
    hipStream_t s = S(stream);
    const int64_t nf = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    k_adj_project_faces<<<cdiv(nf, 256), 256, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    k_adj_project_cells<<<cdiv((int64_t)nx * ny * nz, 256), 256, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_adj_solve_beta(const float* p, const float* lam, float* const g_beta[3], int nx, int ny, int nz,
                             pf_stream_t stream) {
    PF_REQUIRE(nx > 0 && ny > 0 && nz > 0, "pf_adj_solve_beta: bad grid");
    PF_REQUIRE(p && lam && g_beta && g_beta[0] && g_beta[1] && g_beta[2], "pf_adj_solve_beta: null argument");
    const int64_t nf = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    k_adj_solve_beta<<<cdiv(nf, 256), 256, 0, S(stream)>>>(Dims{nx, ny, nz}, p, lam, g_beta[0], g_beta[1],
                                                            g_beta[2]);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_adj_phase(const pf_phase_args* fwd, const float* const g_vel[3], const float* const g_vel_star[3],
                        const float* const g_beta[3], const float* g_rhs, float* const g_mass[3],
                        float* const g_mom[3], pf_stream_t stream) {
    PF_REQUIRE(fwd, "pf_adj_phase: null args");
    const pf_phase_args& in = *fwd;
    PF_REQUIRE(in.nx > 0 && in.ny > 0 && in.nz > 0 && in.dt != 0.0f, "pf_adj_phase: bad grid or dt");
    PF_REQUIRE(!(in.k_div > 0.0f), "pf_adj_phase: k_div > 0 has no adjoint here");
    PF_REQUIRE(in.is_liquid && g_rhs && g_vel && g_vel_star && g_beta && g_mass && g_mom, "pf_adj_phase: null argument");
    AdjPhaseK a;
    a.d = Dims{in.nx, in.ny, in.nz};
    for (int q = 0; q < 3; ++q) {
        PF_REQUIRE(in.mass[q] && in.mom[q] && in.beta[q], "pf_adj_phase: need mass, mom, beta");
        PF_REQUIRE(g_vel[q] && g_vel_star[q] && g_beta[q] && g_mass[q] && g_mom[q], "pf_adj_phase: null argument");
        a.mass[q] = in.mass[q];
        a.mom[q] = in.mom[q];
        a.beta[q] = in.beta[q];
        a.g_vel2[q] = g_vel[q];
        a.g_vel_star[q] = g_vel_star[q];
        a.g_beta[q] = g_beta[q];
        a.g_mass[q] = g_mass[q];
        a.g_mom[q] = g_mom[q];
    }
    a.is_liq = in.is_liquid;
    a.solid = in.solid;
    a.g_rhs = g_rhs;
    a.mass_eps = in.mass_eps;
    a.rho_min = in.rho_min;
    a.alpha_phi = in.alpha_phi;
    a.rho_0l = in.rho_0l;
    a.rho_l = in.rho_l;
    a.rho_g = in.rho_g;
    a.dt = in.dt;
    a.dx = in.dx;
    const int64_t nf = (int64_t)(in.nx + 1) * (in.ny + 1) * (in.nz + 1);
    k_adj_phase<<<cdiv(nf, 256), 256, 0, S(stream)>>>(a);
    PF_CHECK_LAUNCH();
    return PF_OK;
}

PF_API int pf_adj_p2g(const float* px, const float* pv, const float* pmass, const int32_t* pstate, int64_t n,
                      const float* const g_mass[3], const float* const g_mom[3], int nx, int ny, int nz, float dx,
                      float rp, float* gx, float* gv, pf_stream_t stream) {
    PF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && dx > 0.0f && n >= 0, "pf_adj_p2g: bad sizes");
    PF_REQUIRE(rp == dx, "pf_adj_p2g: r_p == dx is required (2 candidates per axis)");
    PF_REQUIRE(px && pv && pmass && pstate && g_mass && g_mom && gx && gv, "pf_adj_p2g: null argument");
    AdjP2GK a;
    a.px = px;
    a.pv = pv;
    a.pmass = pmass;
    a.pstate = pstate;
    a.n = n;
    a.d = Dims{nx, ny, nz};
    a.dx = dx;
    a.inv_r2 = 1.0f / (rp * rp);
    for (int q = 0; q < 3; ++q) {
        PF_REQUIRE(g_mass[q] && g_mom[q], "pf_adj_p2g: null argument");
        a.g_mass[q] = g_mass[q];
        a.g_mom[q] = g_mom[q];
    }
    a.gx = gx;
    a.gv = gv;
    if (n == 0) return PF_OK;
    k_adj_p2g<<<cdiv(n, 256), 256, 0, S(stream)>>>(a);
    PF_CHECK_LAUNCH();
    return PF_OK;
}
