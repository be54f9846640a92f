// nrq_map.h -- block-to-work mapping shared by the lattice kernels.
//
// A launch is a 1-D grid of ngroup * nsb blocks (modes 3 and 4: nsb *
// ceil(ngroup / wpb)): nsb site blocks times `ngroup` work items per site
// block (the three link directions of the leapfrog, the field chunks of the
// CN stencil), the items of one site block adjacent.  Blocks are dealt round-robin to the XCDs, each with its own L2;
// with xcd > 1 the hardware block id is remapped so that each XCD walks one
// contiguous range of site blocks, and the neighbour data a block needs was
// fetched by the same XCD shortly before.
//
// Site blocks are either `bs` consecutive sites (mode 0) or z-sweep tiles
// (mode 1): tile t = yband * n2 + z holds ty full x-rows at one z, so the
// z +- 1 neighbours of a tile are the tiles just before and after it.
//
// Mode 3 (wave items): a site block is 64 consecutive sites, one per lane,
// and each wavefront of a block takes another work item of the same 64
// sites (up to wpb items per block, more blocks per site block beyond that).
// The items of a site read the same links, and in one block those loads are
// served by the CU's L1 instead of L2.
//
// Mode 4 (banded wave items): as mode 3, but the sites are ordered by (y-band
// of ty rows, z, site within the band-plane): a band-plane, ty x-rows at one
// z, is a contiguous run of ty * n0 sites, cut into 64-site blocks (only the
// last may be partial).  With the XCD remap each XCD walks whole bands, one
// z-plane of a band after the other: the y neighbours of a site are in the
// same band-plane (except at the band's edges) and its z neighbours in the
// band-planes just before and after it, so the neighbour values a block reads
// were fetched into the same L2 a few band-planes earlier.
#pragma once
#include "nrq_su3.h"

namespace nrq {

struct lf_map {
    int64_t nsb; // site blocks (modes 0, 3, 4) or tiles (modes 1, 2)
    int bs;      // threads per block
    int xcd;     // XCD remap factor (1: none)
    int mode;    // 0: bs consecutive sites; 1: z-sweep tiles of full x-rows; 2: z-sweep tx x ty tiles;
                 // 3: wave items over 64 consecutive sites; 4: wave items over 64-site blocks of band-planes
    int ty;      // modes 1, 2: y rows per tile; mode 4: y rows per band
    int ngroup;  // work items per site block
    int tx;      // mode 2: x extent of a tile
    int nxb;     // mode 2: tiles along x; mode 4: 64-site blocks per band-plane
    int wpb;     // modes 3, 4: work items (wavefronts) per block
};

// Blocks of a launch.
__host__ __device__ __forceinline__ unsigned lf_blocks(const lf_map &m)
{
    if (m.mode >= 3) return (unsigned)m.nsb * (unsigned)((m.ngroup + m.wpb - 1) / m.wpb);
    return (unsigned)m.ngroup * (unsigned)m.nsb;
}

// Logical block of hardware block blockIdx.x among `total` when XCD x
// (= blockIdx.x mod xcd) is to walk the x-th contiguous range of logical
// blocks.
__device__ __forceinline__ unsigned xcd_block(unsigned total, int xcdi)
{
    const unsigned b = blockIdx.x;
    if (xcdi <= 1) return b;
    const unsigned xcd = (unsigned)xcdi;
    const unsigned q = total / xcd, r = total % xcd;
    const unsigned x = b % xcd, k = b / xcd;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ unsigned lf_logical_block(const lf_map &m)
{
    return xcd_block(lf_blocks(m), m.xcd);
}

// Site of this thread, the work item (direction / chunk) of its block and the
// block's site-block index t (0 <= t < nsb).
__device__ __forceinline__ bool lf_site(const lf_map &m, const geom &g, lsite &p, int &item, int64_t &t)
{
    const unsigned lb = lf_logical_block(m);
    if (m.mode >= 3) {
        const unsigned nbg = (unsigned)((m.ngroup + m.wpb - 1) / m.wpb); // blocks per site block
        const unsigned tt = lb / nbg;
        item = (int)(lb - tt * nbg) * m.wpb + (int)(threadIdx.x >> 6);
        t = tt;
        if (m.mode == 3) {
            const int64_t s = (int64_t)tt * 64 + (threadIdx.x & 63);
            if (s >= g.vol || item >= m.ngroup) return false;
            p = make_site(g, s);
            return true;
        }
        const unsigned per = (unsigned)m.nxb;     // blocks per band-plane
        const unsigned pb = per * (unsigned)g.n2; // blocks per band
        const unsigned band = tt / pb, r1 = tt - band * pb;
        const unsigned z = r1 / per, seg = r1 - z * per;
        const unsigned o = seg * 64u + (threadIdx.x & 63u); // site within the band-plane
        if (o >= (unsigned)m.ty * (unsigned)g.n0 || item >= m.ngroup) return false;
        const unsigned row = o / (unsigned)g.n0;
        const int x = (int)(o - row * (unsigned)g.n0), y = (int)(band * (unsigned)m.ty + row);
        p = lsite{x + (int64_t)g.n0 * (y + (int64_t)g.n1 * (int)z), x, y, (int)z};
        return true;
    }
    const unsigned ng = (unsigned)m.ngroup;
    const unsigned tt = lb / ng;
    item = (int)(lb - tt * ng);
    t = tt;
    if (m.mode == 0) {
        const int64_t s = (int64_t)tt * m.bs + threadIdx.x;
        if (s >= g.vol) return false;
        p = make_site(g, s);
        return true;
    }
    const unsigned n2 = (unsigned)g.n2;
    const unsigned col = tt / n2;
    const int z = (int)(tt - col * n2);
    if (m.mode == 1) {
        const int x = (int)threadIdx.x % g.n0, yy = (int)threadIdx.x / g.n0;
        const int y = (int)col * m.ty + yy;
        if (yy >= m.ty || y >= g.n1) return false;
        p = lsite{x + (int64_t)g.n0 * (y + (int64_t)g.n1 * z), x, y, z};
        return true;
    }
    const unsigned yb = col / (unsigned)m.nxb, xb = col - yb * (unsigned)m.nxb;
    const int x = (int)xb * m.tx + (int)threadIdx.x % m.tx, y = (int)yb * m.ty + (int)threadIdx.x / m.tx;
    if (x >= g.n0 || y >= g.n1) return false;
    p = lsite{x + (int64_t)g.n0 * (y + (int64_t)g.n1 * z), x, y, z};
    return true;
}

__device__ __forceinline__ bool lf_site(const lf_map &m, const geom &g, lsite &p, int &item)
{
    int64_t t;
    return lf_site(m, g, p, item, t);
}

// Host: map for `bs` threads per block.  Mode 1 needs n0 <= bs; mode 2
// uses 16-wide tiles (bs a multiple of 16); modes 3 and 4 take up to bs / 64
// work items per block (bs is then 64 times the items per block), mode 4
// with bands of the largest of 8, 4, 2, 1 rows that divides n1; otherwise
// consecutive sites.
inline lf_map make_lf_map(const geom &g, int bs, int xcd, int mode, int ngroup)
{
    lf_map m{};
    m.bs = bs;
    m.xcd = xcd;
    m.ngroup = ngroup;
    if ((mode == 3 || mode == 4) && bs >= 64) {
        m.mode = mode;
        m.wpb = ngroup < bs / 64 ? ngroup : bs / 64;
        m.bs = 64 * m.wpb;
        if (mode == 3) {
            m.nsb = (g.vol + 63) / 64;
        } else {
            m.ty = g.n1 % 8 == 0 ? 8 : g.n1 % 4 == 0 ? 4 : g.n1 % 2 == 0 ? 2 : 1;
            m.nxb = (m.ty * g.n0 + 63) / 64;
            m.nsb = (int64_t)m.nxb * (g.n1 / m.ty) * g.n2;
        }
    } else if (mode == 1 && g.n0 <= bs) {
        m.mode = 1;
        m.ty = bs / g.n0 < g.n1 ? bs / g.n0 : g.n1;
        m.nsb = (int64_t)((g.n1 + m.ty - 1) / m.ty) * g.n2;
    } else if (mode == 2 && bs % 16 == 0 && g.n0 >= 16) {
        m.mode = 2;
        m.tx = 16;
        m.ty = bs / 16;
        m.nxb = (g.n0 + 15) / 16;
        m.nsb = (int64_t)m.nxb * ((g.n1 + m.ty - 1) / m.ty) * g.n2;
    } else {
        m.mode = 0;
        m.nsb = (g.vol + bs - 1) / bs;
    }
    return m;
}

} // namespace nrq
