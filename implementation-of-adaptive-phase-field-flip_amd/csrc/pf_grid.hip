// Grid-side kernels: fused eq. (7) phase field / beta / walls / gravity / RHS,
// pressure projection + divergence, liquid-velocity extrapolation.
#include "pf_common.hpp"

namespace pf {

int fill_u32(void* dst, int64_t n, uint32_t v, hipStream_t s);

// ============================================================== phase / beta / forces
struct PhaseK {
    Dims d;
    int flags;
    const float* mass[3];
    const float* mom[3];
    float* vel[3];
    float* vel_star[3];
    int* valid[3];
    const float* cmass;
    const float* cwsum;
    float* phi_cell;
    int* is_liq;
    float* phi_face[3];
    float* beta[3];
    const int* solid;
    float* rhs;
    float mass_eps, rho_min, alpha_phi, rho_0l, rho_l, rho_g;
    float g[3];
    float gdt[3];  // g_a * dt, formed once so every pass adds the same value
    float dt, dx, rho_t0, k_div;
};

// eq. (7), kernels.py:234-239
__device__ inline float phase_field(float rho, const PhaseK& a) {
    if (rho < a.rho_min) return 0.0f;
    return fminf(sqrtf(fmaxf(rho - a.rho_min, 0.0f) / (a.alpha_phi * a.rho_0l)), 1.0f);
}

// Liquid tag of a cell: eq. (7) phase >= 1/2 from its mass, or the is_liquid
// entry with PF_PH_TAGS_IN.
__device__ inline int cell_tag(const PhaseK& a, int fl, float cmass, int tag) {
    if (fl & PF_PH_TAGS_IN) return tag;
    return phase_field(cmass, a) >= 0.5f ? 1 : 0;
}

// The inputs of one face: mass (with PF_PH_VEL_FROM_MOM, or with
// PF_PH_PHASE unless VEL_ONLY) and momentum (PF_PH_VEL_FROM_MOM) or velocity.
struct FaceIn {
    float m, u;
};

template <int A, bool VEL_ONLY>
__device__ inline FaceIn load_face(const PhaseK& a, int fl, int f) {
    FaceIn r{0.0f, 0.0f};
    const int need_m = VEL_ONLY ? PF_PH_VEL_FROM_MOM : (PF_PH_VEL_FROM_MOM | PF_PH_PHASE);
    if (fl & need_m) r.m = a.mass[A][f];
    r.u = (fl & PF_PH_VEL_FROM_MOM) ? a.mom[A][f] : a.vel[A][f];
    return r;
}

// face_velocity_from_momentum (kernels.py:212-228), or the stored velocity
__device__ inline float face_vel(const PhaseK& a, int fl, const FaceIn& in, bool& ok) {
    if (fl & PF_PH_VEL_FROM_MOM) {
        ok = in.m > a.mass_eps;
        return ok ? in.u / in.m : 0.0f;
    }
    ok = true;
    return in.u;
}

// Velocity of a cell's high face after every velocity stage selected by
// flags, without writing anything (the cell's RHS needs it; the point that
// owns the face writes it).  The face's low cell is the cell itself (solid
// flag s_lo); hi_out: the face is the domain's high wall; sol: solids given.
template <int A>
__device__ inline float high_face_vel(const PhaseK& a, int fl, bool sol, const FaceIn& in, bool hi_out,
                                      int s_lo, int s_hi) {
    bool ok;
    float v = face_vel(a, fl, in, ok);
    // walls (kernels.py:314-326) and solid-touching faces (kernels.py:601-627)
    if ((fl & PF_PH_ZERO) && (hi_out || (sol && (s_lo == 1 || s_hi == 1)))) return 0.0f;
    if ((fl & PF_PH_GRAVITY) && a.g[A] != 0.0f) v += a.gdt[A];
    return v;
}

// Low face f of axis A: all selected stages, with writes; returns the final
// velocity (for the owning cell's RHS).  lo / hi: the face is the domain's
// low / high wall; t_lo, t_hi: liquid tags of its two cells (used only when
// the cell exists); blocked: zeroed by PF_PH_ZERO.
template <int A>
__device__ inline float low_face(const PhaseK& a, int fl, int f, bool lo, bool hi, const FaceIn& in,
                                 int t_lo, int t_hi, bool blocked) {
    bool ok;
    float v = face_vel(a, fl, in, ok);
    if ((fl & PF_PH_VEL_FROM_MOM) && a.valid[A]) a.valid[A][f] = ok ? 1 : 0;
    bool write_beta = false;
    float beta = 0.0f;
    if (fl & PF_PH_PHASE) {
        // face_beta, kernels.py:278-311
        int ta = lo ? 0 : t_lo;
        int tb = hi ? 0 : t_hi;
        if (lo) ta = tb;
        if (hi) tb = ta;
        float ph;
        if (ta == 1 && tb == 1)
            ph = 1.0f;
        else if (ta == 0 && tb == 0)
            ph = 0.0f;
        else
            ph = phase_field(in.m, a);
        a.phi_face[A][f] = ph;
        const float rho = (1.0f - ph) * a.rho_g + ph * a.rho_l;
        beta = 1.0f / rho;
        write_beta = true;
    }
    if (blocked) {
        v = 0.0f;
        beta = 0.0f;
        write_beta = true;
    }
    if (fl & PF_PH_STAR) a.vel_star[A][f] = v;
    if (fl & PF_PH_GRAVITY) {
        if (a.g[A] != 0.0f) v += a.gdt[A];
        if (blocked) v = 0.0f;
    }
    if (fl & (PF_PH_VEL_FROM_MOM | PF_PH_ZERO | PF_PH_GRAVITY)) a.vel[A][f] = v;
    if (write_beta) a.beta[A][f] = beta;
    return v;
}

// One thread per point of the (nx+1)(ny+1)(nz+1) lattice: cell (i,j,k) and
// its three low faces; the last layer of each axis only has faces.  Every
// input of a point is loaded first, from clamped indices (each addresses an
// existing entry and is the true index wherever its value is used), and the
// values are pinned before the first branch on them: the point's loads are
// in flight together instead of one round trip per dependent branch.
// FL >= 0: the flags (and kPhaseSolid for a solid array) fixed at compile
// time, for the step's combinations; FL < 0: a.flags at run time.
constexpr int kPhaseSolid = 1 << 8;
template <int FL>
__global__ __launch_bounds__(256) void k_phase(PhaseK a) {
    const Dims d = a.d;
    // 32-bit decode (the entry points check that the lattice has fewer
    // than 2^31 points)
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned ez = d.nz + 1, ey = d.ny + 1;
    if (t >= (unsigned)(d.nx + 1) * ey * ez) return;
    const unsigned row = t / ez;
    const int k = (int)(t - row * ez), i = (int)(row / ey), j = (int)(row - (unsigned)i * ey);
    const bool ci = i < d.nx, cj = j < d.ny, ck = k < d.nz;
    const int fl = FL >= 0 ? FL : a.flags;
    const bool sol = FL >= 0 ? (FL & kPhaseSolid) != 0 : a.solid != nullptr;
    const bool rhs = fl & PF_PH_RHS;
    const int ic = min(i, d.nx - 1), jc = min(j, d.ny - 1), kc = min(k, d.nz - 1);
    const int c = cell_index(d, ic, jc, kc);
    // low neighbours (the low cells of the low faces), high neighbours (the
    // high cells of the high faces)
    const int cn0 = cell_index(d, max(i - 1, 0), jc, kc), cn1 = cell_index(d, ic, max(j - 1, 0), kc),
              cn2 = cell_index(d, ic, jc, max(k - 1, 0));
    const int cp0 = cell_index(d, min(i + 1, d.nx - 1), jc, kc),
              cp1 = cell_index(d, ic, min(j + 1, d.ny - 1), kc),
              cp2 = cell_index(d, ic, jc, min(k + 1, d.nz - 1));
    const int f0 = face_index<0>(d, i, jc, kc), f1 = face_index<1>(d, ic, j, kc),
              f2 = face_index<2>(d, ic, jc, k);

    const FaceIn l0 = load_face<0, false>(a, fl, f0), l1 = load_face<1, false>(a, fl, f1),
                 l2 = load_face<2, false>(a, fl, f2);
    float cm = 0.f, cm0 = 0.f, cm1 = 0.f, cm2 = 0.f;
    int tg = 0, tg0 = 0, tg1 = 0, tg2 = 0;
    if (fl & PF_PH_PHASE) {
        cm = a.cmass[c];
        if (fl & PF_PH_TAGS_IN) {
            tg = a.is_liq[c];
            tg0 = a.is_liq[cn0];
            tg1 = a.is_liq[cn1];
            tg2 = a.is_liq[cn2];
        } else {
            cm0 = a.cmass[cn0];
            cm1 = a.cmass[cn1];
            cm2 = a.cmass[cn2];
        }
    }
    int sc = 0, sn0 = 0, sn1 = 0, sn2 = 0, sp0 = 0, sp1 = 0, sp2 = 0;
    if ((fl & PF_PH_ZERO) && sol) {
        sc = a.solid[c];
        sn0 = a.solid[cn0];
        sn1 = a.solid[cn1];
        sn2 = a.solid[cn2];
        if (rhs) {
            sp0 = a.solid[cp0];
            sp1 = a.solid[cp1];
            sp2 = a.solid[cp2];
        }
    }
    FaceIn h0{0.f, 0.f}, h1{0.f, 0.f}, h2{0.f, 0.f};
    float ws = 0.f;
    if (rhs) {
        h0 = load_face<0, true>(a, fl, face_index<0>(d, ic + 1, jc, kc));
        h1 = load_face<1, true>(a, fl, face_index<1>(d, ic, jc + 1, kc));
        h2 = load_face<2, true>(a, fl, face_index<2>(d, ic, jc, kc + 1));
        if (a.k_div > 0.0f) ws = a.cwsum[c];
    }
    // one statement: the compiler may move loads across separate ones
    asm volatile("" ::"v"(l0.m), "v"(l0.u), "v"(l1.m), "v"(l1.u), "v"(l2.m), "v"(l2.u), "v"(cm), "v"(cm0),
                 "v"(cm1), "v"(cm2), "v"(tg), "v"(tg0), "v"(tg1), "v"(tg2), "v"(sc), "v"(sn0), "v"(sn1),
                 "v"(sn2), "v"(sp0), "v"(sp1), "v"(sp2), "v"(h0.m), "v"(h0.u), "v"(h1.m), "v"(h1.u),
                 "v"(h2.m), "v"(h2.u), "v"(ws));

    const int t_c = cell_tag(a, fl, cm, tg);
    const bool zero = fl & PF_PH_ZERO;
    float ulo = 0.f, vlo = 0.f, wlo = 0.f;
    if (cj && ck) {
        const bool lo = i == 0, hi = i == d.nx;
        ulo = low_face<0>(a, fl, f0, lo, hi, l0, cell_tag(a, fl, cm0, tg0), t_c,
                          zero && (lo || hi || (sol && (sn0 == 1 || sc == 1))));
    }
    if (ci && ck) {
        const bool lo = j == 0, hi = j == d.ny;
        vlo = low_face<1>(a, fl, f1, lo, hi, l1, cell_tag(a, fl, cm1, tg1), t_c,
                          zero && (lo || hi || (sol && (sn1 == 1 || sc == 1))));
    }
    if (ci && cj) {
        const bool lo = k == 0, hi = k == d.nz;
        wlo = low_face<2>(a, fl, f2, lo, hi, l2, cell_tag(a, fl, cm2, tg2), t_c,
                          zero && (lo || hi || (sol && (sn2 == 1 || sc == 1))));
    }
    if (!(ci && cj && ck)) return;
    if (fl & PF_PH_PHASE) {
        // cell_phase, kernels.py:242-255
        const float ph = phase_field(cm, a);
        a.phi_cell[c] = ph;
        if (!(fl & PF_PH_TAGS_IN)) a.is_liq[c] = ph >= 0.5f ? 1 : 0;
    }
    if (rhs) {
        // build_rhs, kernels.py:338-363
        const float uhi = high_face_vel<0>(a, fl, sol, h0, i + 1 == d.nx, sc, sp0);
        const float vhi = high_face_vel<1>(a, fl, sol, h1, j + 1 == d.ny, sc, sp1);
        const float whi = high_face_vel<2>(a, fl, sol, h2, k + 1 == d.nz, sc, sp2);
        const float du = (uhi - ulo) + (vhi - vlo) + (whi - wlo);
        float r = -(a.dx / a.dt) * du;
        if (a.k_div > 0.0f) {
            const float excess = (ws - a.rho_t0) / a.rho_t0;
            r -= a.dx * a.dx * a.k_div * fmaxf(excess, 0.0f) / a.dt;
        }
        a.rhs[c] = r;
    }
}

// ============================================================== projection
struct ProjK {
    Dims d;
    const float* p;
    const float* beta[3];
    float* vel[3];
    const float* vel_star[3];
    float* dvel[3];
    float dx, dt;
};

// Face f of axis A with its loaded inputs: v (vel), b (beta), ph / pl
// (pressure of the high / low cell, used only inside the domain), vs
// (vel_star, with dvel); ia / na: the face's index and count along A.
template <int A>
__device__ inline void project_face(const ProjK& a, int f, int ia, int na, float v, float b, float ph, float pl,
                                    float vs) {
    if (a.dt != 0.0f) {
        // kernels.py:376-387: skip beta == 0 and the two boundary planes
        if (b != 0.0f && ia > 0 && ia < na) {
            v -= a.dt * b * (ph - pl) / a.dx;
            a.vel[A][f] = v;
        }
    }
    if (a.dvel[A]) a.dvel[A][f] = v - vs;
}

// One thread per lattice point: its three low faces.  Every input is loaded
// first from clamped indices (valid addresses, the true index wherever the
// value is used) and pinned before the branches, so a point's loads travel
// together.
__global__ __launch_bounds__(256) void k_project(ProjK a) {
    const Dims d = a.d;
    // 32-bit decode (the entry points check that the lattice has fewer
    // than 2^31 points)
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned ez = d.nz + 1, ey = d.ny + 1;
    if (t >= (unsigned)(d.nx + 1) * ey * ez) return;
    const unsigned row = t / ez;
    const int k = (int)(t - row * ez), i = (int)(row / ey), j = (int)(row - (unsigned)i * ey);
    const bool ci = i < d.nx, cj = j < d.ny, ck = k < d.nz;
    const int ic = min(i, d.nx - 1), jc = min(j, d.ny - 1), kc = min(k, d.nz - 1);
    const int f0 = face_index<0>(d, i, jc, kc), f1 = face_index<1>(d, ic, j, kc),
              f2 = face_index<2>(d, ic, jc, k);
    const float v0 = a.vel[0][f0], v1 = a.vel[1][f1], v2 = a.vel[2][f2];
    float b0 = 0.f, b1 = 0.f, b2 = 0.f, pc = 0.f, p0 = 0.f, p1 = 0.f, p2 = 0.f;
    if (a.dt != 0.0f) {
        b0 = a.beta[0][f0];
        b1 = a.beta[1][f1];
        b2 = a.beta[2][f2];
        pc = a.p[cell_index(d, ic, jc, kc)];
        p0 = a.p[cell_index(d, max(i - 1, 0), jc, kc)];
        p1 = a.p[cell_index(d, ic, max(j - 1, 0), kc)];
        p2 = a.p[cell_index(d, ic, jc, max(k - 1, 0))];
    }
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    if (a.dvel[0]) s0 = a.vel_star[0][f0];
    if (a.dvel[1]) s1 = a.vel_star[1][f1];
    if (a.dvel[2]) s2 = a.vel_star[2][f2];
    asm volatile("" ::"v"(v0), "v"(v1), "v"(v2), "v"(b0), "v"(b1), "v"(b2), "v"(pc), "v"(p0), "v"(p1), "v"(p2),
                 "v"(s0), "v"(s1), "v"(s2));
    if (cj && ck) project_face<0>(a, f0, i, d.nx, v0, b0, pc, p0, s0);
    if (ci && ck) project_face<1>(a, f1, j, d.ny, v1, b1, pc, p1, s1);
    if (ci && cj) project_face<2>(a, f2, k, d.nz, v2, b2, pc, p2, s2);
}

// max |div u / dx| (kernels.py:390-400); out must start at 0.  Few wide
// workgroups, one atomic each: at 128^3, 2048 same-address atomics from
// 256-thread workgroups took most of its 27 us.
constexpr int kMaxDivThreads = 1024;
__global__ __launch_bounds__(kMaxDivThreads) void k_max_div(Dims d, const float* __restrict__ u,
                                                            const float* __restrict__ v,
                                                            const float* __restrict__ w, float dx, float* out) {
    __shared__ float sh[kMaxDivThreads / 64];
    float m = 0.0f;
    const unsigned nc = (unsigned)d.nx * d.ny * d.nz;
    for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const unsigned row = c / d.nz;
        const int k = (int)(c - row * d.nz), i = (int)(row / d.ny), j = (int)(row - (unsigned)i * d.ny);
        const float dv = (u[face_index<0>(d, i + 1, j, k)] - u[face_index<0>(d, i, j, k)]) +
                         (v[face_index<1>(d, i, j + 1, k)] - v[face_index<1>(d, i, j, k)]) +
                         (w[face_index<2>(d, i, j, k + 1)] - w[face_index<2>(d, i, j, k)]);
        m = fmaxf(m, fabsf(dv / dx));
    }
    m = block_max(m, sh);
    if (threadIdx.x == 0) atomic_max_nonneg(out, m);
}

// ============================================================== extrapolation
struct ExtrapK {
    Dims d;
    int nb0, nb1;  // workgroups of the x and y face arrays (decode_face)
    const float* vin[3];
    const uint8_t* kin[3];
    float* vout[3];
    uint8_t* kout[3];
    const int* is_liq;
};

// The face this thread owns: the workgroups cover the three face arrays one
// after another (nb0 workgroups for the x faces, nb1 for the y faces, the
// rest for the z faces), so the array is the same in a whole workgroup and
// its pointers stay scalar.  32-bit indices: the caller checks that the
// arrays hold fewer than 2^31 faces.
__device__ inline bool decode_face(const Dims& d, int nb0, int nb1, int& A, int& i, int& j, int& k, int& f) {
    const int b = blockIdx.x;
    A = b < nb0 ? 0 : (b < nb0 + nb1 ? 1 : 2);
    const int t = (b - (A == 0 ? 0 : (A == 1 ? nb0 : nb0 + nb1))) * blockDim.x + threadIdx.x;
    const unsigned sx = d.nx + (A == 0), sy = d.ny + (A == 1), sz = d.nz + (A == 2);
    if ((unsigned)t >= sx * sy * sz) return false;
    const unsigned row = (unsigned)t / sz;
    f = t;
    k = (int)((unsigned)t - row * sz);
    i = (int)(row / sy);
    j = (int)(row - (unsigned)i * sy);
    return true;
}

// seed_liquid_velocity, kernels.py:406-425
__global__ __launch_bounds__(256) void k_extrap_seed(ExtrapK a) {
    int A, i, j, k, f;
    if (!decode_face(a.d, a.nb0, a.nb1, A, i, j, k, f)) return;
    const Dims d = a.d;
    // the two cells' tags and the face value loaded at once (clamped
    // indices, masked), not the value after the tags
    const int im = i - (A == 0), jm = j - (A == 1), km = k - (A == 2);
    const bool hi = i < d.nx && j < d.ny && k < d.nz, lo = im >= 0 && jm >= 0 && km >= 0;
    const int th = a.is_liq[cell_index(d, min(i, d.nx - 1), min(j, d.ny - 1), min(k, d.nz - 1))];
    const int tl = a.is_liq[cell_index(d, max(im, 0), max(jm, 0), max(km, 0))];
    const float v = a.vin[A][f];
    asm volatile("" ::"v"(th), "v"(tl), "v"(v));
    int liq = 0;
    if (hi) liq = max(liq, th);
    if (lo) liq = max(liq, tl);
    a.kout[A][f] = (uint8_t)liq;
    a.vout[A][f] = liq == 1 ? v : 0.0f;
}

// extrapolate_sweep, kernels.py:428-462.  Every load is issued at once (the
// face's own value and flag, and its six neighbours' from indices clamped
// into the array, masked as unknown where the neighbour does not exist):
// loads that waited for the face's flag and then for the neighbours' flags
// made three dependent round trips per face.  The known neighbours are
// summed in the reference's order.
__global__ __launch_bounds__(256) void k_extrap_sweep(ExtrapK a) {
    int A, i, j, k, f;
    if (!decode_face(a.d, a.nb0, a.nb1, A, i, j, k, f)) return;
    const float* vin = a.vin[A];
    const uint8_t* kin = a.kin[A];
    const Dims d = a.d;
    const int sx = d.nx + (A == 0), sy = d.ny + (A == 1), sz = d.nz + (A == 2);
    const uint8_t kc = kin[f];
    const float vc = vin[f];
    uint8_t nk[6];
    float nv[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const int ax = q / 2, s = (q % 2 == 0) ? -1 : 1;
        const int ii = i + (ax == 0 ? s : 0), jj = j + (ax == 1 ? s : 0), kk = k + (ax == 2 ? s : 0);
        const bool in = ii >= 0 && jj >= 0 && kk >= 0 && ii < sx && jj < sy && kk < sz;
        // (clamped: a missing neighbour reads the face itself, masked out)
        const int g = (min(max(ii, 0), sx - 1) * sy + min(max(jj, 0), sy - 1)) * sz + min(max(kk, 0), sz - 1);
        nk[q] = kin[g];
        nv[q] = vin[g];
        if (!in) nk[q] = 0;
    }
    // an empty asm that takes every loaded value keeps the compiler from
    // sinking the neighbours' loads into the branches that use them (which
    // costs one round trip per branch): all 14 loads are in flight together
#pragma unroll
    for (int q = 0; q < 6; ++q) asm volatile("" ::"v"((int)nk[q]), "v"(nv[q]));
    asm volatile("" ::"v"((int)kc), "v"(vc));
    if (kc == 1) {
        a.vout[A][f] = vc;
        a.kout[A][f] = 1;
        return;
    }
    float acc = 0.0f;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q)
        if (nk[q] == 1) {
            acc += nv[q];
            cnt += 1;
        }
    if (cnt > 0) {
        a.vout[A][f] = acc / (float)cnt;
        a.kout[A][f] = 1;
    } else {
        a.vout[A][f] = vc;
        a.kout[A][f] = 0;
    }
}

}  // namespace pf

using namespace pf;

// The step's flag combinations get kernels with the flags fixed at compile
// time (no branches between a point's loads); the others read them.
static void launch_phase(const PhaseK& a, unsigned nb, hipStream_t s) {
    constexpr int kStep = PF_PH_VEL_FROM_MOM | PF_PH_PHASE | PF_PH_ZERO | PF_PH_STAR | PF_PH_GRAVITY | PF_PH_RHS;
    const int key = a.flags | (a.solid ? kPhaseSolid : 0);
    switch (key) {
        case kStep: k_phase<kStep><<<nb, 256, 0, s>>>(a); break;
        case kStep | kPhaseSolid: k_phase<kStep | kPhaseSolid><<<nb, 256, 0, s>>>(a); break;
        case kStep | PF_PH_TAGS_IN: k_phase<kStep | PF_PH_TAGS_IN><<<nb, 256, 0, s>>>(a); break;
        case kStep | PF_PH_TAGS_IN | kPhaseSolid:
            k_phase<kStep | PF_PH_TAGS_IN | kPhaseSolid><<<nb, 256, 0, s>>>(a);
            break;
        default: k_phase<-1><<<nb, 256, 0, s>>>(a);
    }
}

PF_API int pf_phase_beta_forces(const pf_phase_args* args, pf_stream_t stream) {
    PF_REQUIRE(args, "pf_phase_beta_forces: null args");
    const pf_phase_args& in = *args;
    PF_REQUIRE(in.nx > 0 && in.ny > 0 && in.nz > 0, "pf_phase_beta_forces: bad grid");
    PF_REQUIRE((int64_t)(in.nx + 1) * (in.ny + 1) * (in.nz + 1) < (int64_t)1 << 31,
               "pf_phase_beta_forces: grid too large");
    const int fl = in.flags;
    PF_REQUIRE(fl != 0 && (fl & ~127) == 0, "pf_phase_beta_forces: bad flags %d", fl);
    PF_REQUIRE(!(fl & PF_PH_TAGS_IN) || ((fl & PF_PH_PHASE) && in.is_liquid),
               "pf_phase_beta_forces: PF_PH_TAGS_IN goes with PF_PH_PHASE and an is_liquid input");
    // a cell's RHS reads its high faces, which the neighbouring points write
    // in the same pass when vel is an input that ZERO or GRAVITY modifies
    PF_REQUIRE(!((fl & PF_PH_RHS) && !(fl & PF_PH_VEL_FROM_MOM) && (fl & (PF_PH_ZERO | PF_PH_GRAVITY))),
               "pf_phase_beta_forces: PF_PH_RHS with PF_PH_ZERO or PF_PH_GRAVITY needs "
               "PF_PH_VEL_FROM_MOM (or a separate PF_PH_RHS pass)");
    for (int q = 0; q < 3; ++q) {
        if (fl & PF_PH_VEL_FROM_MOM)
            PF_REQUIRE(in.mass[q] && in.mom[q], "pf_phase_beta_forces: need mass/mom");
        PF_REQUIRE(in.vel[q], "pf_phase_beta_forces: need vel");
        if (fl & PF_PH_PHASE)
            PF_REQUIRE(in.mass[q] && in.phi_face[q] && in.beta[q], "pf_phase_beta_forces: phase needs mass, phi_face, beta");
        if (fl & (PF_PH_ZERO | PF_PH_GRAVITY)) PF_REQUIRE(in.beta[q], "pf_phase_beta_forces: need beta");
        if (fl & PF_PH_STAR) PF_REQUIRE(in.vel_star[q], "pf_phase_beta_forces: need vel_star");
    }
    if (fl & PF_PH_PHASE)
        PF_REQUIRE(in.cell_mass && in.phi_cell && in.is_liquid, "pf_phase_beta_forces: phase needs cell arrays");
    if (fl & PF_PH_RHS) {
        PF_REQUIRE(in.rhs, "pf_phase_beta_forces: need rhs");
        PF_REQUIRE(in.dt != 0.0f, "pf_phase_beta_forces: rhs needs dt != 0");
        PF_REQUIRE(!(in.k_div > 0.0f) || in.cell_wsum, "pf_phase_beta_forces: k_div needs cell_wsum");
    }
    PhaseK a;
    a.d = Dims{in.nx, in.ny, in.nz};
    a.flags = fl;
    for (int q = 0; q < 3; ++q) {
        a.mass[q] = in.mass[q];
        a.mom[q] = in.mom[q];
        a.vel[q] = in.vel[q];
        a.vel_star[q] = in.vel_star[q];
        a.valid[q] = in.valid[q];
        a.phi_face[q] = in.phi_face[q];
        a.beta[q] = in.beta[q];
        a.g[q] = in.gravity[q];
        a.gdt[q] = in.gravity[q] * in.dt;
    }
    a.cmass = in.cell_mass;
    a.cwsum = in.cell_wsum;
    a.phi_cell = in.phi_cell;
    a.is_liq = in.is_liquid;
    a.solid = in.solid;
    a.rhs = in.rhs;
    a.mass_eps = in.mass_eps;
    a.rho_min = in.rho_min;
    a.alpha_phi = in.alpha_phi;
    a.rho_0l = in.rho_0l;
    a.rho_l = in.rho_l;
    a.rho_g = in.rho_g;
    a.dt = in.dt;
    a.dx = in.dx;
    a.rho_t0 = in.rho_tilde_0;
    a.k_div = in.k_div;
    const int64_t n = (int64_t)(in.nx + 1) * (in.ny + 1) * (in.nz + 1);
    launch_phase(a, cdiv(n, 256), S(stream));
    PF_CHECK_LAUNCH();
    return PF_OK;
}

// build_rhs alone: the grid pass above with only its rhs stage, which reads
// vel and writes nothing but rhs
PF_API int pf_build_rhs(int32_t nx, int32_t ny, int32_t nz, const float* const vel[3], const float* cell_wsum,
                        float* rhs, float dx, float dt, float k_div, float rho_tilde_0, pf_stream_t stream) {
    PF_REQUIRE(vel, "pf_build_rhs: null vel");
    pf_phase_args a{};
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.flags = PF_PH_RHS;
    for (int q = 0; q < 3; ++q) a.vel[q] = const_cast<float*>(vel[q]);
    a.cell_wsum = cell_wsum;
    a.rhs = rhs;
    a.dx = dx;
    a.dt = dt;
    a.k_div = k_div;
    a.rho_tilde_0 = rho_tilde_0;
    return pf_phase_beta_forces(&a, stream);
}

PF_API int pf_project(const float* p, float* const beta[3], float* const vel[3],
                      float* const vel_star[3], float* const dvel[3], int nx, int ny, int nz,
                      float dx, float dt, float* max_div, pf_stream_t stream) {
    PF_REQUIRE(nx > 0 && ny > 0 && nz > 0, "pf_project: bad grid");
    PF_REQUIRE(vel && vel[0] && vel[1] && vel[2], "pf_project: need vel");
    PF_REQUIRE(dt == 0.0f || (p && beta && beta[0] && beta[1] && beta[2]), "pf_project: need p, beta");
    PF_REQUIRE(dx > 0.f, "pf_project: dx must be positive");
    PF_REQUIRE((int64_t)(nx + 1) * (ny + 1) * (nz + 1) < (int64_t)1 << 31, "pf_project: grid too large");
    hipStream_t s = S(stream);
    const Dims d{nx, ny, nz};
    ProjK a;
    a.d = d;
    a.p = p;
    bool want_dvel = dvel && vel_star;
    for (int q = 0; q < 3; ++q) {
        a.beta[q] = beta ? beta[q] : nullptr;
        a.vel[q] = vel[q];
        a.vel_star[q] = want_dvel ? vel_star[q] : nullptr;
        a.dvel[q] = want_dvel ? dvel[q] : nullptr;
        if (want_dvel) PF_REQUIRE(vel_star[q] && dvel[q], "pf_project: partial dvel/vel_star");
    }
    a.dx = dx;
    a.dt = dt;
    const int64_t n = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    if (dt != 0.0f || want_dvel) {
        k_project<<<cdiv(n, 256), 256, 0, s>>>(a);
        PF_CHECK_LAUNCH();
    }
    if (max_div) {
        int rc = fill_u32(max_div, 1, 0u, s);
        if (rc) return rc;
        const int64_t nb = cdiv((int64_t)nx * ny * nz, kMaxDivThreads * 4);
        k_max_div<<<(unsigned)(nb < 512 ? nb : 512), kMaxDivThreads, 0, s>>>(d, vel[0], vel[1], vel[2], dx,
                                                                             max_div);
        PF_CHECK_LAUNCH();
    }
    return PF_OK;
}

PF_API int pf_extrapolate(float* const vel[3], const int32_t* is_liquid, float* const vel_liq[3],
                          float* const tmp[3], uint8_t* const known[3], uint8_t* const known2[3],
                          int nx, int ny, int nz, int sweeps, pf_stream_t stream) {
    PF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && sweeps >= 0, "pf_extrapolate: bad sizes");
    PF_REQUIRE(vel && is_liquid && vel_liq && tmp && known && known2, "pf_extrapolate: null args");
    hipStream_t s = S(stream);
    const Dims d{nx, ny, nz};
    const int64_t n = face_count(d, 0) + face_count(d, 1) + face_count(d, 2);
    PF_REQUIRE(n < ((int64_t)1 << 31), "pf_extrapolate: grid too large");
    // ping-pong so that the last sweep lands in vel_liq
    float* const* bufs[2] = {vel_liq, tmp};
    uint8_t* const* kb[2] = {known, known2};
    int cur = (sweeps % 2 == 0) ? 0 : 1;
    ExtrapK a;
    a.d = d;
    a.is_liq = is_liquid;
    for (int q = 0; q < 3; ++q) {
        a.vin[q] = vel[q];
        a.kin[q] = nullptr;
        a.vout[q] = bufs[cur][q];
        a.kout[q] = kb[cur][q];
    }
    a.nb0 = (int)cdiv(face_count(d, 0), 256);
    a.nb1 = (int)cdiv(face_count(d, 1), 256);
    const unsigned nbs = (unsigned)(a.nb0 + a.nb1) + cdiv(face_count(d, 2), 256);
    k_extrap_seed<<<nbs, 256, 0, s>>>(a);
    PF_CHECK_LAUNCH();
    for (int sw = 0; sw < sweeps; ++sw) {
        const int nxt = 1 - cur;
        for (int q = 0; q < 3; ++q) {
            a.vin[q] = bufs[cur][q];
            a.kin[q] = kb[cur][q];
            a.vout[q] = bufs[nxt][q];
            a.kout[q] = kb[nxt][q];
        }
        k_extrap_sweep<<<nbs, 256, 0, s>>>(a);
        PF_CHECK_LAUNCH();
        cur = nxt;
    }
    return PF_OK;
}
