// MSBG device addressing shared by the Poisson stack and the particle
// kernels (pfflip/msbg.py:81-141): every lookup goes through finest-level
// integer coordinates, block = F >> log2(n0), cell = (F & (n0 - 1)) >> l.
#pragma once

#include "pf_common.hpp"

namespace pf {

// face fields hold 3 entries per cell, indexed in 32-bit arithmetic
constexpr int kMsbgMaxCells = 715827882;  // floor((2^31 - 1) / 3)

struct MGrid {
    int nbx, nby, nbz, n0, sh0, nslots, ncells;
    float h0;
    const int* level;
    const int* slot;
    const int* blk;
    const int* blev;
    const int* coff;
    const int* cslot;
};

struct MCell {
    int addr, slot, level;
};

inline MGrid mgrid(const pf_msbg_grid& g) {
    MGrid m;
    m.nbx = g.nbx;
    m.nby = g.nby;
    m.nbz = g.nbz;
    m.n0 = g.n0;
    m.sh0 = g.sh0;
    m.nslots = g.nslots;
    m.ncells = g.ncells;
    m.h0 = g.h0;
    m.level = g.level;
    m.slot = g.slot;
    m.blk = g.blk;
    m.blev = g.blev;
    m.coff = g.coff;
    m.cslot = g.cslot;
    return m;
}

// msbg.py:81-108
__device__ inline MCell locate(const MGrid& g, int fx, int fy, int fz) {
    MCell c{-1, -1, -1};
    if (fx < 0 || fy < 0 || fz < 0) return c;
    const int bx = fx >> g.sh0, by = fy >> g.sh0, bz = fz >> g.sh0;
    if (bx >= g.nbx || by >= g.nby || bz >= g.nbz) return c;
    const int s = g.slot[(bx * g.nby + by) * g.nbz + bz];
    if (s < 0) return c;
    const int l = g.blev[s];
    const int n = g.n0 >> l, m = g.n0 - 1;
    const int i = (fx & m) >> l, j = (fy & m) >> l, k = (fz & m) >> l;
    c.addr = g.coff[s] + (i * n + j) * n + k;
    c.slot = s;
    c.level = l;
    return c;
}

// finest coordinate of the cell's low corner, and its level (msbg.py:111-141)
__device__ inline int4 origin(const MGrid& g, int addr) {
    const int s = g.cslot[addr];
    const int l = g.blev[s];
    const int n = g.n0 >> l, ln = g.sh0 - l;
    const int r = addr - g.coff[s];
    const int i = r >> (2 * ln), j = (r >> ln) & (n - 1), k = r & (n - 1);
    const int b = g.blk[s];
    const int bz = b % g.nbz, by = (b / g.nbz) % g.nby, bx = b / (g.nbz * g.nby);
    return make_int4((bx << g.sh0) + (i << l), (by << g.sh0) + (j << l), (bz << g.sh0) + (k << l), l);
}

__device__ inline int comp(int axis, int c) { return axis == c ? 1 : 0; }

// A cell's block: slot, level, cells per axis, first address and the
// block's low corner in finest coordinates.
struct BlockRef {
    int slot, level, n, base, bx0, by0, bz0;
};

__device__ inline BlockRef block_ref(const MGrid& g, int s) {
    BlockRef r;
    r.slot = s;
    r.level = g.blev[s];
    r.n = g.n0 >> r.level;
    r.base = g.coff[s];
    const int b = g.blk[s];
    r.bz0 = (b % g.nbz) << g.sh0;
    r.by0 = ((b / g.nbz) % g.nby) << g.sh0;
    r.bx0 = (b / (g.nbz * g.nby)) << g.sh0;
    return r;
}

// locate() for a finest coordinate near block `r`: inside that block the
// address follows from the local indices without the lookup chain
// (slot -> level, offset); elsewhere locate().  Same result as locate().
__device__ inline MCell locate_near(const MGrid& g, const BlockRef& r, int fx, int fy, int fz) {
    const unsigned ux = (unsigned)(fx - r.bx0), uy = (unsigned)(fy - r.by0), uz = (unsigned)(fz - r.bz0);
    if ((ux | uy | uz) < (unsigned)g.n0) {
        MCell c;
        c.addr = r.base + (((int)ux >> r.level) * r.n + ((int)uy >> r.level)) * r.n + ((int)uz >> r.level);
        c.slot = r.slot;
        c.level = r.level;
        return c;
    }
    return locate(g, fx, fy, fz);
}

// origin() and the cell's block in one pass
__device__ inline int4 origin_ref(const MGrid& g, int addr, BlockRef& r) {
    r = block_ref(g, g.cslot[addr]);
    // n = 2^(sh0 - level): the local indices are bit fields of the offset
    const int q = addr - r.base, ln = g.sh0 - r.level;
    const int i = q >> (2 * ln), j = (q >> ln) & (r.n - 1), k = q & (r.n - 1);
    return make_int4(r.bx0 + (i << r.level), r.by0 + (j << r.level), r.bz0 + (k << r.level), r.level);
}

// the cell containing a world position (msbg.py:154-161)
__device__ inline MCell locate_at(const MGrid& g, float x, float y, float z) {
    return locate(g, (int)floorf(x / g.h0), (int)floorf(y / g.h0), (int)floorf(z / g.h0));
}

}  // namespace pf
