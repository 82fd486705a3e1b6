// C++ / OpenMP restatement of the reference's uniform-grid PF-FLIP step, FP32.
//
// TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the
// GPU ("CPU restatement of the reference algorithm", kind "port"), and a
// second restatement that tests/ check against the FP64 numpy oracle.  The
// product library never links or calls it.
//
// It follows the reference kernel by kernel, in the reference's launch
// order and with its data layout (k-fastest grids, per-axis face arrays of
// (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1); AoS float3 particles):
//   p2g_face / p2g_cell: per-particle scatter with float atomics
//                              (pfflip/kernels.py:100-209)
//   face_velocity_from_momentum, cell_phase, face_beta, zero_solid_faces,
//   zero_solid_obstacle_faces  (kernels.py:212-326, :601-627)
//   vel_star copy, add_gravity (sim.py:176-178, :270-285; kernels.py:332-335)
//   build_rhs                  (kernels.py:338-363)
//   MGPCG: compute_diag, galerkin_coarsen_faces, rbgs_sweep, residual,
//   restrict_sum, prolong_inject_add, flexible PCG with Polak-Ribiere beta
//                              (pfflip/poisson.py:40-409)
//   apply_pressure_gradient, divergence_norm (kernels.py:366-400)
//   field_difference + g2p     (kernels.py:468-502)
//   seed_liquid_velocity + extrapolate_sweep (kernels.py:406-462, sim.py:329-341)
//   advect_rk3, reflect_at_walls, push_out_of_solids (kernels.py:508-549,567-595,630-679)
// The droplet stage is not restated (the measured configurations run with
// droplets off).  Two arithmetic differences, both shared with the GPU
// solver: the operator row is evaluated as sum_f beta_f (p_i - p_nbr) (see
// apply_row), and reductions (dot products, sums, maxima) accumulate in
// FP64 where the reference sums FP32 atomics into one float; the scalars
// alpha and beta are FP64 on the host in both.  With one thread the scatter
// uses plain adds (a serial launch), with more it uses OpenMP atomics.

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using Vec = std::vector<float>;

struct Dims {
    int nx, ny, nz;
    int64_t n() const { return (int64_t)nx * ny * nz; }
    int64_t at(int i, int j, int k) const { return ((int64_t)i * ny + j) * nz + k; }
};

Dims face_dims(const Dims& d, int axis) {
    return {d.nx + (axis == 0), d.ny + (axis == 1), d.nz + (axis == 2)};
}

struct Params {
    // layout shared with the ctypes wrapper (oracle/cpu_step.py)
    double dx, r_p, rho_l, rho_g, alpha_phi, alpha_flip_liquid, alpha_flip_air;
    double gravity[3];
    double cfl, dt_min, dt_max, pcg_tol;
    double omega, rho_tilde_0, rho_0l, rho_min, k_div;
    int32_t pcg_max_iter, pre, post, coarsest, min_dim, extrap_cells;
};

struct Level {
    Dims d;
    Vec b3[3], diag, p, b, r;
};

struct Sim {
    Dims d;
    Params P;
    Vec vel[3], vel_star[3], vel_liq[3], tmp[3], dvel[3], mass[3], mom[3], beta[3], phi_face[3];
    std::vector<int32_t> valid[3], known[3], known2[3];
    Vec cell_mass, cell_wsum, phi_cell, pressure, rhs;
    std::vector<int32_t> is_liquid, solid;
    bool has_solids = false;
    int64_t np = 0;
    Vec px, pv, pmass, prad;
    std::vector<int32_t> ptype, pstate;
    std::vector<Level> lv;
    Vec cg_rhs, res_v, z, q, Ap, r_prev;
};

inline float splat_weight(float d2, float r2) {
    // eq. (6), kernels.py:40-49
    float t = 1.0f - d2 / r2;
    t = std::max(t, 0.0f);
    return t * t * t;
}

template <bool ATOMIC>
inline void add(float& dst, float v) {
    if (ATOMIC) {
#pragma omp atomic
        dst += v;
    } else {
        dst += v;
    }
}

// kernels.py:100-158 (axis 0..2, face offsets) and :161-209 (axis 3: cell
// centres, weight sum instead of momentum)
template <bool ATOMIC>
void p2g_pass(Sim& s, int axis) {
    const Dims fd = axis < 3 ? face_dims(s.d, axis) : s.d;
    const float dx = (float)s.P.dx;
    const float o[3] = {axis == 0 ? 0.0f : 0.5f, axis == 1 ? 0.0f : 0.5f, axis == 2 ? 0.0f : 0.5f};
    float* m = axis < 3 ? s.mass[axis].data() : s.cell_mass.data();
    float* mu = axis < 3 ? s.mom[axis].data() : s.cell_wsum.data();
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < s.np; ++p) {
        if (s.pstate[p] != 0) continue;
        const float* x = &s.px[3 * p];
        const float mp = s.pmass[p], rp = s.prad[p];
        const float up = axis < 3 ? s.pv[3 * p + axis] : 0.0f;
        const float r2 = rp * rp, R = rp / dx;
        const float cx = x[0] / dx - o[0], cy = x[1] / dx - o[1], cz = x[2] / dx - o[2];
        const int ilo = std::max((int)std::floor(cx - R) + 1, 0), ihi = std::min((int)std::floor(cx + R), fd.nx - 1);
        const int jlo = std::max((int)std::floor(cy - R) + 1, 0), jhi = std::min((int)std::floor(cy + R), fd.ny - 1);
        const int klo = std::max((int)std::floor(cz - R) + 1, 0), khi = std::min((int)std::floor(cz + R), fd.nz - 1);
        for (int i = ilo; i <= ihi; ++i) {
            const float fx = ((float)i + o[0]) * dx - x[0];
            for (int j = jlo; j <= jhi; ++j) {
                const float fy = ((float)j + o[1]) * dx - x[1];
                for (int k = klo; k <= khi; ++k) {
                    const float fz = ((float)k + o[2]) * dx - x[2];
                    const float w = splat_weight(fx * fx + fy * fy + fz * fz, r2);
                    if (w > 0.0f) {
                        const int64_t f = fd.at(i, j, k);
                        add<ATOMIC>(m[f], w * mp);
                        add<ATOMIC>(mu[f], axis < 3 ? w * mp * up : w);
                    }
                }
            }
        }
    }
}

inline float phase_field(float rho, const Params& P) {
    // eq. (7), kernels.py:234-238
    if (rho < (float)P.rho_min) return 0.0f;
    return std::min(std::sqrt(std::max(rho - (float)P.rho_min, 0.0f) / ((float)P.alpha_phi * (float)P.rho_0l)),
                    1.0f);
}

// the two cells of face (i,j,k) of `axis`; -1 outside the grid
inline void face_cells(const Dims& d, int axis, int i, int j, int k, int64_t& lo, int64_t& hi) {
    const int im = i - (axis == 0), jm = j - (axis == 1), km = k - (axis == 2);
    lo = (im < 0 || jm < 0 || km < 0) ? -1 : d.at(im, jm, km);
    hi = (i >= d.nx || j >= d.ny || k >= d.nz) ? -1 : d.at(i, j, k);
}

inline bool wall_face(const Dims& fd, int axis, int i, int j, int k) {
    const int n = axis == 0 ? i : (axis == 1 ? j : k);
    const int last = (axis == 0 ? fd.nx : (axis == 1 ? fd.ny : fd.nz)) - 1;
    return n == 0 || n == last;
}

// zero_solid_faces + zero_solid_obstacle_faces (kernels.py:314-326, :601-627)
void zero_blocked(Sim& s, int axis) {
    const Dims fd = face_dims(s.d, axis);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < fd.nx; ++i)
        for (int j = 0; j < fd.ny; ++j)
            for (int k = 0; k < fd.nz; ++k) {
                const int64_t f = fd.at(i, j, k);
                bool blocked = wall_face(fd, axis, i, j, k);
                if (s.has_solids && !blocked) {
                    int64_t lo, hi;
                    face_cells(s.d, axis, i, j, k, lo, hi);
                    blocked = (lo >= 0 && s.solid[lo] == 1) || (hi >= 0 && s.solid[hi] == 1);
                }
                if (blocked) {
                    s.beta[axis][f] = 0.0f;
                    s.vel[axis][f] = 0.0f;
                }
            }
}

// ------------------------------------------------------------------ P2G .. forces
void p2g(Sim& s) {
    // sim.py:205-236
    for (int a = 0; a < 3; ++a) {
        std::fill(s.mass[a].begin(), s.mass[a].end(), 0.0f);
        std::fill(s.mom[a].begin(), s.mom[a].end(), 0.0f);
    }
    std::fill(s.cell_mass.begin(), s.cell_mass.end(), 0.0f);
    std::fill(s.cell_wsum.begin(), s.cell_wsum.end(), 0.0f);
    const bool par = omp_get_max_threads() > 1;
    for (int a = 0; a < 4; ++a) {
        if (par)
            p2g_pass<true>(s, a);
        else
            p2g_pass<false>(s, a);
    }
    const float eps = (float)(1e-12 * s.P.rho_0l);
    for (int a = 0; a < 3; ++a) {
        const int64_t n = (int64_t)s.mass[a].size();
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < n; ++f) {
            const float m = s.mass[a][f];
            if (m > eps) {
                s.vel[a][f] = s.mom[a][f] / m;
                s.valid[a][f] = 1;
            } else {
                s.vel[a][f] = 0.0f;
                s.valid[a][f] = 0;
            }
        }
    }
}

void phase_and_coefficients(Sim& s) {
    // sim.py:238-268
    const int64_t nc = s.d.n();
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
        const float f = phase_field(s.cell_mass[c], s.P);
        s.phi_cell[c] = f;
        s.is_liquid[c] = f >= 0.5f ? 1 : 0;
    }
    const float rho_l = (float)s.P.rho_l, rho_g = (float)s.P.rho_g;
    for (int a = 0; a < 3; ++a) {
        const Dims fd = face_dims(s.d, a);
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < fd.nx; ++i)
            for (int j = 0; j < fd.ny; ++j)
                for (int k = 0; k < fd.nz; ++k) {
                    // face_beta, kernels.py:258-311
                    int64_t lo, hi;
                    face_cells(s.d, a, i, j, k, lo, hi);
                    int ta = lo >= 0 ? s.is_liquid[lo] : 0;
                    int tb = hi >= 0 ? s.is_liquid[hi] : 0;
                    if (lo < 0) ta = tb;
                    if (hi < 0) tb = ta;
                    const int64_t f = fd.at(i, j, k);
                    float phi;
                    if (ta == 1 && tb == 1)
                        phi = 1.0f;
                    else if (ta == 0 && tb == 0)
                        phi = 0.0f;
                    else
                        phi = phase_field(s.mass[a][f], s.P);
                    s.phi_face[a][f] = phi;
                    s.beta[a][f] = 1.0f / ((1.0f - phi) * rho_g + phi * rho_l);
                }
        zero_blocked(s, a);
    }
}

void add_forces(Sim& s, float dt) {
    // sim.py:270-285
    for (int a = 0; a < 3; ++a) {
        const float g = (float)s.P.gravity[a];
        if (g != 0.0f) {
            const int64_t n = (int64_t)s.vel[a].size();
#pragma omp parallel for schedule(static)
            for (int64_t f = 0; f < n; ++f) s.vel[a][f] += g * dt;
        }
        zero_blocked(s, a);
    }
}

void build_rhs(Sim& s, float dt) {
    // kernels.py:338-363
    const Dims d = s.d;
    const Dims fx = face_dims(d, 0), fy = face_dims(d, 1), fz = face_dims(d, 2);
    const float dx = (float)s.P.dx, rho0 = (float)s.P.rho_tilde_0, kd = (float)s.P.k_div;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j)
            for (int k = 0; k < d.nz; ++k) {
                const float du = (s.vel[0][fx.at(i + 1, j, k)] - s.vel[0][fx.at(i, j, k)]) +
                                 (s.vel[1][fy.at(i, j + 1, k)] - s.vel[1][fy.at(i, j, k)]) +
                                 (s.vel[2][fz.at(i, j, k + 1)] - s.vel[2][fz.at(i, j, k)]);
                float r = -(dx / dt) * du;
                const int64_t c = d.at(i, j, k);
                if (kd > 0.0f) r -= dx * dx * kd * std::max((s.cell_wsum[c] - rho0) / rho0, 0.0f) / dt;
                s.rhs[c] = r;
            }
}

// ------------------------------------------------------------------ MGPCG (poisson.py)
inline float apply_row(const Level& L, const float* p, int i, int j, int k) {
    // _apply_row, poisson.py:54-76 (clamped neighbour reads), written as
    // sum_f beta_f (p_i - p_nbr) rather than diag p_i - sum_f beta_f p_nbr:
    // the same operator, but the reference form's cancellation of two
    // nearly equal FP32 numbers floors the recursive residual near 1e-6 at
    // 1000:1 contrast, above the 1e-6 tolerance of the measured workload
    // (the GPU solver uses the difference form for the same reason)
    const Dims d = L.d;
    const Dims fx = face_dims(d, 0), fy = face_dims(d, 1), fz = face_dims(d, 2);
    const float pc = p[d.at(i, j, k)];
    float s = L.b3[0][fx.at(i, j, k)] * (pc - p[d.at(std::max(i - 1, 0), j, k)]);
    s += L.b3[0][fx.at(i + 1, j, k)] * (pc - p[d.at(std::min(i + 1, d.nx - 1), j, k)]);
    s += L.b3[1][fy.at(i, j, k)] * (pc - p[d.at(i, std::max(j - 1, 0), k)]);
    s += L.b3[1][fy.at(i, j + 1, k)] * (pc - p[d.at(i, std::min(j + 1, d.ny - 1), k)]);
    s += L.b3[2][fz.at(i, j, k)] * (pc - p[d.at(i, j, std::max(k - 1, 0))]);
    s += L.b3[2][fz.at(i, j, k + 1)] * (pc - p[d.at(i, j, std::min(k + 1, d.nz - 1))]);
    return s;
}

void setup_levels(Sim& s) {
    // MGPCG.__init__, poisson.py:249-275
    Dims r = s.d;
    while (true) {
        Level L;
        L.d = r;
        for (int a = 0; a < 3; ++a) L.b3[a].assign(face_dims(r, a).n(), 0.0f);
        L.diag.assign(r.n(), 0.0f);
        L.p.assign(r.n(), 0.0f);
        L.b.assign(r.n(), 0.0f);
        L.r.assign(r.n(), 0.0f);
        s.lv.push_back(std::move(L));
        if (std::min({r.nx, r.ny, r.nz}) / 2 < s.P.min_dim || (r.nx % 2) || (r.ny % 2) || (r.nz % 2)) break;
        r = {r.nx / 2, r.ny / 2, r.nz / 2};
    }
    const int64_t n = s.d.n();
    for (Vec* v : {&s.cg_rhs, &s.res_v, &s.z, &s.q, &s.Ap, &s.r_prev}) v->assign(n, 0.0f);
}

void build_hierarchy(Sim& s) {
    // poisson.py:295-309: level 0 copies beta, each level its diagonal, then
    // coarse faces = sum of the 4 fine faces
    for (int a = 0; a < 3; ++a) s.lv[0].b3[a] = s.beta[a];
    for (size_t l = 0; l < s.lv.size(); ++l) {
        Level& L = s.lv[l];
        const Dims d = L.d;
        const Dims fx = face_dims(d, 0), fy = face_dims(d, 1), fz = face_dims(d, 2);
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < d.nx; ++i)
            for (int j = 0; j < d.ny; ++j)
                for (int k = 0; k < d.nz; ++k)
                    L.diag[d.at(i, j, k)] = L.b3[0][fx.at(i, j, k)] + L.b3[0][fx.at(i + 1, j, k)] +
                                            L.b3[1][fy.at(i, j, k)] + L.b3[1][fy.at(i, j + 1, k)] +
                                            L.b3[2][fz.at(i, j, k)] + L.b3[2][fz.at(i, j, k + 1)];
        if (l + 1 < s.lv.size()) {
            Level& C = s.lv[l + 1];
            for (int a = 0; a < 3; ++a) {
                const Dims ff = face_dims(d, a), cf = face_dims(C.d, a);
                const float* fine = L.b3[a].data();
                float* coarse = C.b3[a].data();
#pragma omp parallel for collapse(2) schedule(static)
                for (int I = 0; I < cf.nx; ++I)
                    for (int J = 0; J < cf.ny; ++J)
                        for (int K = 0; K < cf.nz; ++K) {
                            float acc = 0.0f;
                            for (int u = 0; u < 2; ++u)
                                for (int v = 0; v < 2; ++v) {
                                    int ii = 2 * I, jj = 2 * J, kk = 2 * K;
                                    if (a == 0) {
                                        jj += u;
                                        kk += v;
                                    } else if (a == 1) {
                                        ii += u;
                                        kk += v;
                                    } else {
                                        ii += u;
                                        jj += v;
                                    }
                                    if (ii < ff.nx && jj < ff.ny && kk < ff.nz) acc += fine[ff.at(ii, jj, kk)];
                                }
                            coarse[cf.at(I, J, K)] = acc;
                        }
            }
        }
    }
}

void rbgs(Level& L, int colour, float omega) {
    // rbgs_sweep, poisson.py:105-126: in place, one colour
    const Dims d = L.d;
    float* p = L.p.data();
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j)
            for (int k = ((i + j + colour) & 1); k < d.nz; k += 2) {
                const int64_t c = d.at(i, j, k);
                const float dg = L.diag[c];
                if (dg <= 0.0f) continue;
                const float res = L.b[c] - apply_row(L, p, i, j, k);
                p[c] += omega * res / dg;
            }
}

void smooth(Sim& s, Level& L, int n, bool reverse) {
    const float omega = (float)s.P.omega;
    for (int it = 0; it < n; ++it) {
        rbgs(L, reverse ? 1 : 0, omega);
        rbgs(L, reverse ? 0 : 1, omega);
    }
}

double vsum(const Vec& a) {
    double acc = 0.0;
    const int64_t n = (int64_t)a.size();
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (int64_t i = 0; i < n; ++i) acc += a[i];
    return acc;
}

double vdot(const Vec& a, const Vec& b) {
    double acc = 0.0;
    const int64_t n = (int64_t)a.size();
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (int64_t i = 0; i < n; ++i) acc += (double)a[i] * (double)b[i];
    return acc;
}

double vmaxabs(const Vec& a) {
    double m = 0.0;
    const int64_t n = (int64_t)a.size();
#pragma omp parallel for reduction(max : m) schedule(static)
    for (int64_t i = 0; i < n; ++i) m = std::max(m, (double)std::fabs(a[i]));
    return m;
}

void remove_mean(Vec& a) {
    // _remove_mean, poisson.py:291-292
    const float c = (float)(-vsum(a) / (double)a.size());
    const int64_t n = (int64_t)a.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) a[i] += c;
}

void vcycle(Sim& s, size_t l) {
    // _vcycle, poisson.py:323-343
    Level& L = s.lv[l];
    if (l + 1 == s.lv.size()) {
        smooth(s, L, s.P.coarsest, false);
        remove_mean(L.p);
        return;
    }
    smooth(s, L, s.P.pre, false);
    const Dims d = L.d;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j)
            for (int k = 0; k < d.nz; ++k) L.r[d.at(i, j, k)] = L.b[d.at(i, j, k)] - apply_row(L, L.p.data(), i, j, k);
    Level& C = s.lv[l + 1];
    const Dims c = C.d;
#pragma omp parallel for collapse(2) schedule(static)
    for (int I = 0; I < c.nx; ++I)
        for (int J = 0; J < c.ny; ++J)
            for (int K = 0; K < c.nz; ++K) {
                float acc = 0.0f;
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int e = 0; e < 2; ++e) acc += L.r[d.at(2 * I + a, 2 * J + b, 2 * K + e)];
                C.b[c.at(I, J, K)] = acc;
            }
    std::fill(C.p.begin(), C.p.end(), 0.0f);
    vcycle(s, l + 1);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j)
            for (int k = 0; k < d.nz; ++k) L.p[d.at(i, j, k)] += C.p[c.at(i / 2, j / 2, k / 2)];
    smooth(s, L, s.P.post, true);
}

void precondition(Sim& s, const Vec& r, Vec& zout) {
    // poisson.py:345-352
    Level& L0 = s.lv[0];
    L0.b = r;
    std::fill(L0.p.begin(), L0.p.end(), 0.0f);
    vcycle(s, 0);
    zout = L0.p;
    remove_mean(zout);
}

void apply_A(Sim& s, const Vec& p, Vec& out) {
    const Level& L = s.lv[0];
    const Dims d = L.d;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j)
            for (int k = 0; k < d.nz; ++k) out[d.at(i, j, k)] = apply_row(L, p.data(), i, j, k);
}

void axpy(float a, const Vec& x, Vec& y) {
    const int64_t n = (int64_t)y.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
}

int solve(Sim& s, double tol, int max_iter, double* rel_out) {
    // MGPCG.solve, poisson.py:355-409
    s.cg_rhs = s.rhs;
    remove_mean(s.cg_rhs);
    const double bnorm = vmaxabs(s.cg_rhs);
    Vec& x = s.pressure;
    std::fill(x.begin(), x.end(), 0.0f);
    if (bnorm == 0.0) {
        *rel_out = 0.0;
        return 0;
    }
    s.res_v = s.cg_rhs;
    precondition(s, s.res_v, s.z);
    s.q = s.z;
    double rz = vdot(s.res_v, s.z);
    int it = 0;
    double rel = 1.0;
    for (it = 1; it <= max_iter; ++it) {
        apply_A(s, s.q, s.Ap);
        const double qAq = vdot(s.q, s.Ap);
        if (qAq <= 0.0) break;
        const double alpha = rz / qAq;
        axpy((float)alpha, s.q, x);
        s.r_prev = s.res_v;
        axpy((float)-alpha, s.Ap, s.res_v);
        rel = vmaxabs(s.res_v) / bnorm;
        if (rel < tol) break;
        precondition(s, s.res_v, s.z);
        const double rz_new = vdot(s.res_v, s.z);
        const double num = rz_new - vdot(s.r_prev, s.z);
        const double beta = rz != 0.0 ? std::max(num / rz, 0.0) : 0.0;
        const int64_t n = (int64_t)s.q.size();
        const float fb = (float)beta;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) s.q[i] = s.z[i] + fb * s.q[i];
        rz = rz_new;
    }
    remove_mean(x);
    *rel_out = rel;
    return std::min(it, max_iter);
}

// ------------------------------------------------------------------ projection .. advection
void project(Sim& s, float dt) {
    // apply_pressure_gradient, kernels.py:366-387
    const float dx = (float)s.P.dx;
    for (int a = 0; a < 3; ++a) {
        const Dims fd = face_dims(s.d, a);
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < fd.nx; ++i)
            for (int j = 0; j < fd.ny; ++j)
                for (int k = 0; k < fd.nz; ++k) {
                    const int64_t f = fd.at(i, j, k);
                    const float bc = s.beta[a][f];
                    if (bc == 0.0f) continue;
                    int64_t lo, hi;
                    face_cells(s.d, a, i, j, k, lo, hi);
                    if (lo < 0 || hi < 0) continue;
                    s.vel[a][f] -= dt * bc * (s.pressure[hi] - s.pressure[lo]) / dx;
                }
    }
}

float max_divergence(const Sim& s) {
    // divergence_norm, kernels.py:390-400
    const Dims d = s.d;
    const Dims fx = face_dims(d, 0), fy = face_dims(d, 1), fz = face_dims(d, 2);
    const float dx = (float)s.P.dx;
    float m = 0.0f;
#pragma omp parallel for collapse(2) reduction(max : m) schedule(static)
    for (int i = 0; i < d.nx; ++i)
        for (int j = 0; j < d.ny; ++j)
            for (int k = 0; k < d.nz; ++k) {
                const float dv = (s.vel[0][fx.at(i + 1, j, k)] - s.vel[0][fx.at(i, j, k)]) +
                                 (s.vel[1][fy.at(i, j + 1, k)] - s.vel[1][fy.at(i, j, k)]) +
                                 (s.vel[2][fz.at(i, j, k + 1)] - s.vel[2][fz.at(i, j, k)]);
                m = std::max(m, std::fabs(dv / dx));
            }
    return m;
}

inline float sample_face(const Vec& f, const Dims& fd, const float* x, float dx, int axis) {
    // kernels.py:52-80: trilinear with clamped indices
    const float o0 = axis == 0 ? 0.0f : 0.5f, o1 = axis == 1 ? 0.0f : 0.5f, o2 = axis == 2 ? 0.0f : 0.5f;
    const float g0 = x[0] / dx - o0, g1 = x[1] / dx - o1, g2 = x[2] / dx - o2;
    const int i0 = (int)std::floor(g0), j0 = (int)std::floor(g1), k0 = (int)std::floor(g2);
    const float tx = g0 - (float)i0, ty = g1 - (float)j0, tz = g2 - (float)k0;
    float acc = 0.0f;
    for (int di = 0; di < 2; ++di) {
        const float wi = di == 0 ? 1.0f - tx : tx;
        const int ii = std::min(std::max(i0 + di, 0), fd.nx - 1);
        for (int dj = 0; dj < 2; ++dj) {
            const float wj = dj == 0 ? 1.0f - ty : ty;
            const int jj = std::min(std::max(j0 + dj, 0), fd.ny - 1);
            for (int dk = 0; dk < 2; ++dk) {
                const float wk = dk == 0 ? 1.0f - tz : tz;
                const int kk = std::min(std::max(k0 + dk, 0), fd.nz - 1);
                acc += wi * wj * wk * f[fd.at(ii, jj, kk)];
            }
        }
    }
    return acc;
}

inline void sample_velocity(const Sim& s, const Vec* f, const float* x, float* out) {
    const float dx = (float)s.P.dx;
    for (int a = 0; a < 3; ++a) out[a] = sample_face(f[a], face_dims(s.d, a), x, dx, a);
}

void g2p(Sim& s) {
    // field_difference + g2p, kernels.py:468-502, sim.py:316-327
    for (int a = 0; a < 3; ++a) {
        const int64_t n = (int64_t)s.vel[a].size();
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < n; ++f) s.dvel[a][f] = s.vel[a][f] - s.vel_star[a][f];
    }
    const float al = (float)s.P.alpha_flip_liquid, aa = (float)s.P.alpha_flip_air;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < s.np; ++p) {
        if (s.pstate[p] != 0) continue;
        const float alpha = s.ptype[p] == 1 ? al : aa;
        float delta[3], pic[3];
        sample_velocity(s, s.dvel, &s.px[3 * p], delta);
        sample_velocity(s, s.vel, &s.px[3 * p], pic);
        for (int a = 0; a < 3; ++a) s.pv[3 * p + a] = alpha * (s.pv[3 * p + a] + delta[a]) + (1.0f - alpha) * pic[a];
    }
}

void extrapolate(Sim& s) {
    // sim.py:329-341: seed, then extrap_cells Jacobi sweeps with buffer swaps
    for (int a = 0; a < 3; ++a) {
        const Dims fd = face_dims(s.d, a);
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < fd.nx; ++i)
            for (int j = 0; j < fd.ny; ++j)
                for (int k = 0; k < fd.nz; ++k) {
                    int64_t lo, hi;
                    face_cells(s.d, a, i, j, k, lo, hi);
                    int liq = 0;
                    if (hi >= 0) liq = std::max(liq, s.is_liquid[hi]);
                    if (lo >= 0) liq = std::max(liq, s.is_liquid[lo]);
                    const int64_t f = fd.at(i, j, k);
                    s.known[a][f] = liq;
                    s.vel_liq[a][f] = liq == 1 ? s.vel[a][f] : 0.0f;
                }
    }
    for (int sweep = 0; sweep < s.P.extrap_cells; ++sweep) {
        for (int a = 0; a < 3; ++a) {
            const Dims fd = face_dims(s.d, a);
            const Vec& vin = s.vel_liq[a];
            const std::vector<int32_t>& kin = s.known[a];
            Vec& vout = s.tmp[a];
            std::vector<int32_t>& kout = s.known2[a];
#pragma omp parallel for collapse(2) schedule(static)
            for (int i = 0; i < fd.nx; ++i)
                for (int j = 0; j < fd.ny; ++j)
                    for (int k = 0; k < fd.nz; ++k) {
                        const int64_t f = fd.at(i, j, k);
                        if (kin[f] == 1) {
                            vout[f] = vin[f];
                            kout[f] = 1;
                            continue;
                        }
                        float acc = 0.0f;
                        int cnt = 0;
                        for (int dd = 0; dd < 6; ++dd) {
                            const int ax = dd / 2, sg = (dd % 2 == 0) ? -1 : 1;
                            const int ii = i + (ax == 0 ? sg : 0), jj = j + (ax == 1 ? sg : 0),
                                      kk = k + (ax == 2 ? sg : 0);
                            if (ii < 0 || jj < 0 || kk < 0 || ii >= fd.nx || jj >= fd.ny || kk >= fd.nz) continue;
                            const int64_t g = fd.at(ii, jj, kk);
                            if (kin[g] == 1) {
                                acc += vin[g];
                                ++cnt;
                            }
                        }
                        if (cnt > 0) {
                            vout[f] = acc / (float)cnt;
                            kout[f] = 1;
                        } else {
                            vout[f] = vin[f];
                            kout[f] = 0;
                        }
                    }
            std::swap(s.vel_liq[a], s.tmp[a]);
            std::swap(s.known[a], s.known2[a]);
        }
    }
}

void advect(Sim& s, float dt) {
    // sim.py:408-424: advect_rk3, reflect_at_walls, push_out_of_solids
    const float eps = (float)(1e-4 * s.P.dx);
    const float lo[3] = {eps, eps, eps};
    const float hi[3] = {(float)(s.d.nx * s.P.dx - 1e-4 * s.P.dx), (float)(s.d.ny * s.P.dx - 1e-4 * s.P.dx),
                         (float)(s.d.nz * s.P.dx - 1e-4 * s.P.dx)};
    const float dx = (float)s.P.dx;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < s.np; ++p) {
        float* x = &s.px[3 * p];
        float* v = &s.pv[3 * p];
        if (s.pstate[p] == 0) {
            // kernels.py:508-549
            const Vec* f = s.ptype[p] == 1 ? s.vel_liq : s.vel;
            float k1[3], k2[3], k3[3], y[3];
            sample_velocity(s, f, x, k1);
            for (int a = 0; a < 3; ++a) y[a] = x[a] + 0.5f * dt * k1[a];
            sample_velocity(s, f, y, k2);
            for (int a = 0; a < 3; ++a) y[a] = x[a] + 0.75f * dt * k2[a];
            sample_velocity(s, f, y, k3);
            for (int a = 0; a < 3; ++a) {
                const float xn = x[a] + dt * ((2.0f / 9.0f) * k1[a] + (3.0f / 9.0f) * k2[a] + (4.0f / 9.0f) * k3[a]);
                x[a] = std::min(std::max(xn, lo[a]), hi[a]);
            }
        }
        if (s.pstate[p] == 2) continue;
        // kernels.py:567-595
        for (int a = 0; a < 3; ++a) {
            if (x[a] <= lo[a]) {
                x[a] = lo[a];
                v[a] = std::max(v[a], 0.0f);
            } else if (x[a] >= hi[a]) {
                x[a] = hi[a];
                v[a] = std::min(v[a], 0.0f);
            }
        }
        if (!s.has_solids) continue;
        // kernels.py:630-679
        const int i = std::min(std::max((int)(x[0] / dx), 0), s.d.nx - 1);
        const int j = std::min(std::max((int)(x[1] / dx), 0), s.d.ny - 1);
        const int k = std::min(std::max((int)(x[2] / dx), 0), s.d.nz - 1);
        if (s.solid[s.d.at(i, j, k)] == 0) continue;
        float best = 1.0e9f;
        int bax = 0, bsg = 1;
        for (int dd = 0; dd < 6; ++dd) {
            const int ax = dd / 2, sg = (dd % 2 == 0) ? -1 : 1;
            for (int t = 1; t < 8; ++t) {
                const int ii = i + (ax == 0 ? sg * t : 0), jj = j + (ax == 1 ? sg * t : 0),
                          kk = k + (ax == 2 ? sg * t : 0);
                if (ii < 0 || jj < 0 || kk < 0 || ii >= s.d.nx || jj >= s.d.ny || kk >= s.d.nz) break;
                if (s.solid[s.d.at(ii, jj, kk)] == 0) {
                    if ((float)t < best) {
                        best = (float)t;
                        bax = ax;
                        bsg = sg;
                    }
                    break;
                }
            }
        }
        if (best > 1.0e8f) continue;
        x[bax] += (best + 0.5f) * dx * (float)bsg;
        if ((float)bsg * v[bax] < 0.0f) v[bax] = 0.0f;
    }
}

float particle_max_speed(const Sim& s) {
    // kernels.py:552-564
    float m = 0.0f;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (int64_t p = 0; p < s.np; ++p) {
        if (s.pstate[p] != 0) continue;
        const float* v = &s.pv[3 * p];
        m = std::max(m, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    }
    return m;
}

}  // namespace

extern "C" {

void* cs_create(int nx, int ny, int nz, const Params* P) {
    Sim* s = new Sim();
    s->d = {nx, ny, nz};
    s->P = *P;
    for (int a = 0; a < 3; ++a) {
        const int64_t n = face_dims(s->d, a).n();
        for (Vec* v : {&s->vel[a], &s->vel_star[a], &s->vel_liq[a], &s->tmp[a], &s->dvel[a], &s->mass[a],
                       &s->mom[a], &s->beta[a], &s->phi_face[a]})
            v->assign(n, 0.0f);
        s->valid[a].assign(n, 0);
        s->known[a].assign(n, 0);
        s->known2[a].assign(n, 0);
    }
    const int64_t nc = s->d.n();
    for (Vec* v : {&s->cell_mass, &s->cell_wsum, &s->phi_cell, &s->pressure, &s->rhs}) v->assign(nc, 0.0f);
    s->is_liquid.assign(nc, 0);
    s->solid.assign(nc, 0);
    setup_levels(*s);
    return s;
}

void cs_destroy(void* h) { delete static_cast<Sim*>(h); }

int cs_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

// sim.py:142-155: pmass = rho of the phase, prad = r_p, pstate = 0
void cs_set_particles(void* h, int64_t n, const float* x, const float* v, const int32_t* ptype) {
    Sim& s = *static_cast<Sim*>(h);
    s.np = n;
    s.px.assign(x, x + 3 * n);
    s.pv.assign(v, v + 3 * n);
    s.ptype.assign(ptype, ptype + n);
    s.pstate.assign(n, 0);
    s.pmass.resize(n);
    s.prad.assign(n, (float)s.P.r_p);
    for (int64_t p = 0; p < n; ++p) s.pmass[p] = (float)(ptype[p] == 1 ? s.P.rho_l : s.P.rho_g);
}

void cs_set_solids(void* h, const int32_t* mask) {
    Sim& s = *static_cast<Sim*>(h);
    s.solid.assign(mask, mask + s.d.n());
    s.has_solids = std::any_of(s.solid.begin(), s.solid.end(), [](int32_t v) { return v != 0; });
}

double cs_dt_from_cfl(void* h) {
    // sim.py:158-166
    Sim& s = *static_cast<Sim*>(h);
    const double g = std::sqrt(s.P.gravity[0] * s.P.gravity[0] + s.P.gravity[1] * s.P.gravity[1] +
                               s.P.gravity[2] * s.P.gravity[2]);
    double vmax = std::max({(double)particle_max_speed(s), g > 0 ? std::sqrt(2.0 * g * s.P.dx) : 0.0, 1e-8});
    return std::min(std::max(s.P.cfl * s.P.dx / vmax, s.P.dt_min), s.P.dt_max);
}

// one step (sim.py:168-202, droplets off); out = {dt, pcg_iter, pcg_res, max_div}
void cs_step(void* h, double dt, double* out) {
    Sim& s = *static_cast<Sim*>(h);
    if (dt <= 0.0) dt = cs_dt_from_cfl(h);
    const float fdt = (float)dt;
    p2g(s);
    phase_and_coefficients(s);
    for (int a = 0; a < 3; ++a) s.vel_star[a] = s.vel[a];
    add_forces(s, fdt);
    build_rhs(s, fdt);
    build_hierarchy(s);
    double rel = 0.0;
    const int it = solve(s, s.P.pcg_tol, s.P.pcg_max_iter, &rel);
    project(s, fdt);
    const float div = max_divergence(s);
    g2p(s);
    extrapolate(s);
    advect(s, fdt);
    out[0] = dt;
    out[1] = it;
    out[2] = rel;
    out[3] = div;
}

// copy a field out: 0 px, 1 pv, 2 pressure, 3 phi_cell, 4 is_liquid (int32),
// 5..7 vel[a], 8..10 beta[a], 11 rhs, 12 cell_mass
int cs_get(void* h, int which, void* dst) {
    Sim& s = *static_cast<Sim*>(h);
    auto cp = [&](const void* src, size_t bytes) {
        std::memcpy(dst, src, bytes);
        return 0;
    };
    switch (which) {
        case 0: return cp(s.px.data(), s.px.size() * 4);
        case 1: return cp(s.pv.data(), s.pv.size() * 4);
        case 2: return cp(s.pressure.data(), s.pressure.size() * 4);
        case 3: return cp(s.phi_cell.data(), s.phi_cell.size() * 4);
        case 4: return cp(s.is_liquid.data(), s.is_liquid.size() * 4);
        case 5: case 6: case 7: return cp(s.vel[which - 5].data(), s.vel[which - 5].size() * 4);
        case 8: case 9: case 10: return cp(s.beta[which - 8].data(), s.beta[which - 8].size() * 4);
        case 11: return cp(s.rhs.data(), s.rhs.size() * 4);
        case 12: return cp(s.cell_mass.data(), s.cell_mass.size() * 4);
        default: return -1;
    }
}

}  // extern "C"
