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
#include "pf_msbg.hpp"

namespace pf {

int fill_u32(void* dst, int64_t n, uint32_t v, hipStream_t s);

// sum_f T_f (p_A - p_B) over the faces of cell `addr`, or (want_diag) the
// sum of T_f, with T = beta * min(hA, hB)^2 / ((hA + hB) / 2)
// (mr_operator.py:74-166).  The faces are walked in the reference's order.
__device__ inline float apply_row(const MGrid& g, const float* __restrict__ p, const float* __restrict__ bface,
                                  int addr, bool want_diag) {
    const int4 o = origin(g, addr);
    const int la = o.w;
    const int ext = 1 << la;
    const float hA = g.h0 * (float)ext;
    const float pa = want_diag ? 0.0f : p[addr];
    float acc = 0.0f, diag = 0.0f;
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
            const MCell nb = locate(g, fx, fy, fz);
            if (nb.addr < 0) continue;  // domain wall: no flux
            const float hB = g.h0 * (float)(1 << nb.level);
            if (nb.level >= la) {
                // same size or coarser: one neighbour
                float area = fminf(hA, hB);
                area = area * area;
                const float dist = 0.5f * (hA + hB);
                const float beta = side == 0 ? bface[addr * 3 + axis] : bface[nb.addr * 3 + axis];
                const float t = beta * area / dist;
                diag += t;
                if (!want_diag) acc += t * (pa - p[nb.addr]);
            } else {
                // finer: the 2x2 patch across this face
                const int sub = 1 << nb.level;
                const float area = hB * hB;
                const float dist = 0.5f * (hA + hB);
                const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
                for (int u = 0; u < 2; ++u)
                    for (int v = 0; v < 2; ++v) {
                        const int du = u * sub, dv = v * sub;
                        const MCell nc = locate(g, fx + comp(a1, 0) * du + comp(a2, 0) * dv,
                                                fy + comp(a1, 1) * du + comp(a2, 1) * dv,
                                                fz + comp(a1, 2) * du + comp(a2, 2) * dv);
                        if (nc.addr < 0) continue;
                        const float beta = side == 0 ? bface[addr * 3 + axis] : bface[nc.addr * 3 + axis];
                        const float t = beta * area / dist;
                        diag += t;
                        if (!want_diag) acc += t * (pa - p[nc.addr]);
                    }
            }
        }
    }
    return want_diag ? diag : acc;
}

__global__ __launch_bounds__(256) void k_msbg_apply(MGrid g, const float* __restrict__ p,
                                                    const float* __restrict__ bface, float* __restrict__ out,
                                                    int want_diag) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    out[a] = apply_row(g, p, bface, a, want_diag != 0);
}

__global__ __launch_bounds__(256) void k_msbg_residual(MGrid g, const float* __restrict__ p,
                                                       const float* __restrict__ b,
                                                       const float* __restrict__ bface, float* __restrict__ r) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    r[a] = b[a] - apply_row(g, p, bface, a, false);
}

// weighted Gauss-Seidel half-sweep of one colour (mr_operator.py:193-212),
// colour from the finest-coordinate checkerboard, pin -> pout
__global__ __launch_bounds__(256) void k_msbg_rbgs(MGrid g, const float* __restrict__ pin, float* __restrict__ pout,
                                                   const float* __restrict__ b, const float* __restrict__ bface,
                                                   const float* __restrict__ diag, int colour, float omega) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= g.ncells) return;
    const int4 o = origin(g, a);
    float v = pin[a];
    if ((((o.x >> o.w) + (o.y >> o.w) + (o.z >> o.w)) & 1) == colour) {
        const float d = diag[a];
        if (d > 0.0f) v = v + omega * (b[a] - apply_row(g, pin, bface, a, false)) / d;
    }
    pout[a] = v;
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

__global__ __launch_bounds__(64) void k_vec_comb(int op, const double* __restrict__ part, int np,
                                                 double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double r = 0.0;
    for (int b = 0; b < np; ++b) r = op == 2 ? fmax(r, part[b]) : r + part[b];
    *out = r;
}

// y = a x + b y (x may be NULL: y = b y + a)
__global__ __launch_bounds__(256) void k_vec_axpby(float a, const float* __restrict__ x, float b, float* __restrict__ y,
                                                   int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = x ? a * x[i] + b * y[i] : b * y[i] + a;
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
    k_msbg_rbgs<<<cdiv(g->ncells, 256), 256, 0, S(stream)>>>(mgrid(*g), pin, pout, b, bface, diag, colour, omega);
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

PF_API int pf_vec_jacobi(const float* r, const float* diag, float* z, int64_t n, pf_stream_t stream) {
    PF_REQUIRE(r && diag && z && n > 0, "pf_vec_jacobi: null or empty vector");
    const int64_t nb64 = cdiv(n, 256 * 4);
    const int nb = (int)(nb64 < kVecBlocks ? nb64 : kVecBlocks);
    k_vec_jacobi<<<nb, 256, 0, S(stream)>>>(r, diag, z, n);
    PF_CHECK_LAUNCH();
    return PF_OK;
}
