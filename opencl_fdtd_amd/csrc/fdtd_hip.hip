// MI355X (gfx950) kernels and C-ABI for the FDTD time-stepping hot path.
//
// Design: a memory-bound 3-D stencil.  The reference layout is kept
// (component-major, C order, z contiguous; kernels/common.cl:34-61), so the
// (y, z) plane of one x index is contiguous in memory.  A thread owns one
// plane cell p = y*sz + z and marches along x (the slowest axis) through a
// chunk of planes, keeping the previous/next plane values it needs in
// registers; 64 consecutive p of a wavefront give fully coalesced 256-byte
// loads for every component whatever sz is.  In-plane neighbours (p - sz,
// p - 1, ...) are re-read through the L1/L2 caches.  No MFMA: the stencil
// does a few flops per 4-byte word.
//
// Arithmetic follows the reference templates operation by operation
// (kernels/update_e.cl:39-107, kernels/update_h.cl:36-154,
// kernels/update_s.cl:33-36); the file is compiled with -ffp-contract=off so
// every product and sum is rounded separately, as in the CPU restatement.
//
// Poynting S is fused into the H kernel: S at a cell needs the oS products of
// its -x, -y and -z neighbours; the -x ones come from the previous plane of
// the march, the in-plane ones are recomputed from the unmodified H^{n-1/2}
// buffer (H is ping-ponged in this mode, so no block ever reads a value
// another block has already overwritten).  The reference's two-pass dataflow
// (6N oS buffer + S pass) is kept as FDTD_POYNTING_TWO_PASS.

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <string>

#include "../../include/fdtd.h"

namespace {

constexpr int kBlock = 256;          // 4 wave64 per workgroup
constexpr int kTargetBlocks = 4096;  // enough workgroups to fill 256 CUs several times over

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// ---------------------------------------------------------------------------
// Kernel argument blocks (passed by value)
// ---------------------------------------------------------------------------

// One CPML face.  Faces are held per axis in up to two slots, in the order
// they appear in the face list; thick == 0 marks an empty slot.
template <class T>
struct FaceDev {
    int pol, thick;
    int lo, hi;          // slab coordinates [lo, hi) along the axis
    int psi_off;         // offset of psi_u inside a psi block; psi_v follows
    int face_elems;      // elements of psi_u
    int coef_off;        // offset of p0e inside the coefficient block
};

template <class T>
struct Geom {
    int sx, sy, sz, plane;
    int xchunk;
    int seam_mask;       // Bloch axes with the phase applied (0 when decoupled)
    T dt;
    T inv_u;
    const T* inv_e[3];   // nonuniform 1/dx at E locations (H step), or null
    const T* inv_h[3];   // nonuniform 1/dx at H locations (E step), or null
    const T* coef;
    FaceDev<T> face[3][2];   // [axis][slot]
};

template <class T>
struct ESys {
    T* E[3];
    const T* H[3];
    const T* P[3];       // partner of H across a Bloch seam
    T* psi;
    T sa[3], sb[3];      // seam neighbour = sa*own + sb*partner
};

template <class T>
struct HSys {
    const T* E[3];
    const T* Hin[3];
    T* Hout[3];
    const T* P[3];       // partner of E across a Bloch seam
    const T* psi_in;
    T* psi_out;
    T sa[3], sb[3];
};

template <class T>
struct EArgs {
    Geom<T> g;
    const T* eps[3];
    ESys<T> s[2];
    // fused point source: E_comp[src_i] += *src_amp after the update
    const T* src_amp;
    int src_x, src_p, src_comp;
};

template <class T>
struct HArgs {
    Geom<T> g;
    HSys<T> s[2];
    T* S[3];
    T* oS[6];            // two-pass mode: oSxy, oSyz, oSzx, oSxz, oSyx, oSzy
};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

template <class T>
__device__ __forceinline__ bool slab(const FaceDev<T>& f, int c, int s, int* ir) {
    if (c < f.lo || c >= f.hi) return false;
    *ir = (f.pol == 0) ? c : (s - 1) - c;
    return true;
}

// psi index: C order over the grid shape with the face axis replaced by its
// thickness, coordinate ir along that axis.
template <int R, class T>
__device__ __forceinline__ size_t psi_index(const Geom<T>& g, const FaceDev<T>& f, int x, int y, int z, int p,
                                            int ir) {
    if (R == 0) return (size_t)ir * g.plane + p;
    if (R == 1) return ((size_t)x * f.thick + ir) * g.sz + z;
    return ((size_t)x * g.sy + y) * f.thick + ir;
}

// CPML of one face on the derivatives of one cell; update_e.cl:64-100 (ISE)
// and update_h.cl:80-116.  d_ab = d(field_a)/d(b).
template <int R, class T, bool ISE, bool WRITE>
__device__ __forceinline__ void apply_face(const Geom<T>& g, const FaceDev<T>& f, int ir, size_t ip,
                                           const T* psi_in, T* psi_out,
                                           T& dxy, T& dxz, T& dyx, T& dyz, T& dzx, T& dzy,
                                           T& px, T& py, T& pz) {
    const T* c = g.coef + f.coef_off + (ISE ? 0 : 3 * f.thick);
    const T p0 = c[ir], p1 = c[f.thick + ir], p2 = c[2 * f.thick + ir];
    const T* iu = psi_in + f.psi_off;
    const T* iv = iu + f.face_elems;
    T su, sv;
    if (R == 0) {                      // r = x, (u, v) = (y, z)
        dzx *= p2; dyx *= p2;
        su = p0 * iu[ip] + p1 * dzx;
        sv = p0 * iv[ip] + p1 * dyx;
        if (ISE) { py -= su; pz += sv; } else { py += su; pz -= sv; }
    } else if (R == 1) {               // r = y, (u, v) = (x, z)
        dzy *= p2; dxy *= p2;
        su = p0 * iu[ip] + p1 * dzy;
        sv = p0 * iv[ip] + p1 * dxy;
        if (ISE) { px += su; pz -= sv; } else { px -= su; pz += sv; }
    } else {                           // r = z, (u, v) = (x, y)
        dyz *= p2; dxz *= p2;
        su = p0 * iu[ip] + p1 * dyz;
        sv = p0 * iv[ip] + p1 * dxz;
        if (ISE) { px -= su; py += sv; } else { px += su; py -= sv; }
    }
    if (WRITE) {
        T* ou = psi_out + f.psi_off;
        ou[ip] = su;
        ou[ip + f.face_elems] = sv;
    }
}

template <class T>
__device__ __forceinline__ T seam_mix(T own, T partner, T a, T b) {
    return a * own + b * partner;
}

// All CPML faces on one cell: axis x, then y, then z; within an axis the
// faces keep their list order.  Faces on different axes scale different
// derivatives and add at most one term each to a psi-sum that starts at 0,
// so this order gives bit-identical results to any interleaving of the list
// unless two slabs of the same axis overlap (a grid thinner than both).
template <class T, bool ISE, bool WRITE>
__device__ __forceinline__ void cpml_cell(const Geom<T>& g, int x, int y, int z, int p, const T* psi_in,
                                          T* psi_out, T& dxy, T& dxz, T& dyx, T& dyz, T& dzx, T& dzy,
                                          T& px, T& py, T& pz) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const FaceDev<T>& f = g.face[0][k];
        int ir;
        if (f.thick && slab(f, x, g.sx, &ir))
            apply_face<0, T, ISE, WRITE>(g, f, ir, psi_index<0>(g, f, x, y, z, p, ir), psi_in, psi_out,
                                         dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const FaceDev<T>& f = g.face[1][k];
        int ir;
        if (f.thick && slab(f, y, g.sy, &ir))
            apply_face<1, T, ISE, WRITE>(g, f, ir, psi_index<1>(g, f, x, y, z, p, ir), psi_in, psi_out,
                                         dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const FaceDev<T>& f = g.face[2][k];
        int ir;
        if (f.thick && slab(f, z, g.sz, &ir))
            apply_face<2, T, ISE, WRITE>(g, f, ir, psi_index<2>(g, f, x, y, z, p, ir), psi_in, psi_out,
                                         dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    }
}

// ---------------------------------------------------------------------------
// E half-step, update_e.cl:20-107.  One thread per plane cell, marching x.
// ---------------------------------------------------------------------------

template <class T, bool SEAM>
__device__ __forceinline__ void e_cell(const EArgs<T>& a, const ESys<T>& s, int x, int y, int z, int p,
                                       int pym, int pzm, size_t b, T hy_xm, T hz_xm, T& hy, T& hz,
                                       const T* eps0, const T* eps1, const T* eps2, bool src_here) {
    const Geom<T>& g = a.g;
    const T hx = s.H[0][b + p];
    hy = s.H[1][b + p];
    hz = s.H[2][b + p];
    T hx_ym = s.H[0][b + pym], hz_ym = s.H[2][b + pym];
    T hx_zm = s.H[0][b + pzm], hy_zm = s.H[1][b + pzm];
    if (SEAM) {
        if ((g.seam_mask & 2) && y == 0) {
            hx_ym = seam_mix(hx_ym, s.P[0][b + pym], s.sa[1], s.sb[1]);
            hz_ym = seam_mix(hz_ym, s.P[2][b + pym], s.sa[1], s.sb[1]);
        }
        if ((g.seam_mask & 4) && z == 0) {
            hx_zm = seam_mix(hx_zm, s.P[0][b + pzm], s.sa[2], s.sb[2]);
            hy_zm = seam_mix(hy_zm, s.P[1][b + pzm], s.sa[2], s.sb[2]);
        }
    }
    T ix = g.inv_u, iy = g.inv_u, iz = g.inv_u;
    if (g.inv_h[0]) { ix = g.inv_h[0][x]; iy = g.inv_h[1][y]; iz = g.inv_h[2][z]; }
    T dxy = (hx - hx_ym) * iy;
    T dxz = (hx - hx_zm) * iz;
    T dyx = (hy - hy_xm) * ix;
    T dyz = (hy - hy_zm) * iz;
    T dzx = (hz - hz_xm) * ix;
    T dzy = (hz - hz_ym) * iy;
    T px = (T)0, py = (T)0, pz = (T)0;
    cpml_cell<T, true, true>(g, x, y, z, p, s.psi, s.psi, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    const size_t i = b + p;
    T ex = s.E[0][i] + a.g.dt / eps0[i] * (dzy - dyz + px);
    T ey = s.E[1][i] + a.g.dt / eps1[i] * (dxz - dzx + py);
    T ez = s.E[2][i] + a.g.dt / eps2[i] * (dyx - dxy + pz);
    if (src_here) {
        const T amp = *a.src_amp;
        if (a.src_comp == 0) ex = ex + amp;
        else if (a.src_comp == 1) ey = ey + amp;
        else ez = ez + amp;
    }
    s.E[0][i] = ex;
    s.E[1][i] = ey;
    s.E[2][i] = ez;
}

template <class T, bool PAIR, bool SEAM>
__global__ __launch_bounds__(kBlock) void e_march(EArgs<T> a) {
    const Geom<T>& g = a.g;
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= g.plane) return;
    const int y = p / g.sz;
    const int z = p - y * g.sz;
    const int pym = (y == 0) ? p + (g.sy - 1) * g.sz : p - g.sz;
    const int pzm = (z == 0) ? p + g.sz - 1 : p - 1;
    const int x0 = blockIdx.y * g.xchunk;
    const int x1 = min(x0 + g.xchunk, g.sx);
    const int xm = (x0 == 0) ? g.sx - 1 : x0 - 1;
    const size_t bm = (size_t)xm * g.plane + p;
    constexpr int NS = PAIR ? 2 : 1;
    T hy_m[NS], hz_m[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        hy_m[k] = a.s[k].H[1][bm];
        hz_m[k] = a.s[k].H[2][bm];
        if (SEAM && (g.seam_mask & 1) && x0 == 0) {
            hy_m[k] = seam_mix(hy_m[k], a.s[k].P[1][bm], a.s[k].sa[0], a.s[k].sb[0]);
            hz_m[k] = seam_mix(hz_m[k], a.s[k].P[2][bm], a.s[k].sa[0], a.s[k].sb[0]);
        }
    }
    for (int x = x0; x < x1; ++x) {
        const size_t b = (size_t)x * g.plane;
        const bool src_here = a.src_amp != nullptr && x == a.src_x && p == a.src_p;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            T hy, hz;
            e_cell<T, SEAM>(a, a.s[k], x, y, z, p, pym, pzm, b, hy_m[k], hz_m[k], hy, hz,
                            a.eps[0], a.eps[1], a.eps[2], k == 0 && src_here);
            hy_m[k] = hy;
            hz_m[k] = hz;
        }
    }
}

// ---------------------------------------------------------------------------
// H half-step, update_h.cl:22-154, with S fused (update_s.cl:33-36).
// ---------------------------------------------------------------------------

template <class T>
struct HCell {
    T hx0, hy0, hz0, hx1, hy1, hz1;
    T aExy, aExz, aEyx, aEyz, aEzx, aEzy;
};

// The H update of cell (x, y, z).  ex, ey, ez: E at the cell; ey_x1, ez_x1:
// E at the +x neighbour (already mixed across a Bloch x seam).
template <class T, bool SEAM, bool WRITE>
__device__ __forceinline__ HCell<T> h_cell(const Geom<T>& g, const HSys<T>& s, int x, int y, int z, int p,
                                           size_t b, T ex, T ey, T ez, T ey_x1, T ez_x1) {
    const int pyp = (y == g.sy - 1) ? p - (g.sy - 1) * g.sz : p + g.sz;
    const int pzp = (z == g.sz - 1) ? p - (g.sz - 1) : p + 1;
    T ex_y = s.E[0][b + pyp], ez_y = s.E[2][b + pyp];
    T ex_z = s.E[0][b + pzp], ey_z = s.E[1][b + pzp];
    if (SEAM) {
        if ((g.seam_mask & 2) && y == g.sy - 1) {
            ex_y = seam_mix(ex_y, s.P[0][b + pyp], s.sa[1], s.sb[1]);
            ez_y = seam_mix(ez_y, s.P[2][b + pyp], s.sa[1], s.sb[1]);
        }
        if ((g.seam_mask & 4) && z == g.sz - 1) {
            ex_z = seam_mix(ex_z, s.P[0][b + pzp], s.sa[2], s.sb[2]);
            ey_z = seam_mix(ey_z, s.P[1][b + pzp], s.sa[2], s.sb[2]);
        }
    }
    T ix = g.inv_u, iy = g.inv_u, iz = g.inv_u;
    if (g.inv_e[0]) { ix = g.inv_e[0][x]; iy = g.inv_e[1][y]; iz = g.inv_e[2][z]; }
    T dxy = (ex_y - ex) * iy;
    T dxz = (ex_z - ex) * iz;
    T dyx = (ey_x1 - ey) * ix;
    T dyz = (ey_z - ey) * iz;
    T dzx = (ez_x1 - ez) * ix;
    T dzy = (ez_y - ez) * iy;
    T px = (T)0, py = (T)0, pz = (T)0;
    cpml_cell<T, false, WRITE>(g, x, y, z, p, s.psi_in, s.psi_out, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    HCell<T> r;
    const size_t i = b + p;
    r.hx0 = s.Hin[0][i];
    r.hy0 = s.Hin[1][i];
    r.hz0 = s.Hin[2][i];
    r.hx1 = r.hx0 - g.dt * (dzy - dyz - px);
    r.hy1 = r.hy0 - g.dt * (dxz - dzx - py);
    r.hz1 = r.hz0 - g.dt * (dyx - dxy - pz);
    if (WRITE) {
        s.Hout[0][i] = r.hx1;
        s.Hout[1][i] = r.hy1;
        s.Hout[2][i] = r.hz1;
    }
    r.aExy = ex_y + ex; r.aExz = ex_z + ex;
    r.aEyx = ey_x1 + ey; r.aEyz = ey_z + ey;
    r.aEzx = ez_x1 + ez; r.aEzy = ez_y + ez;
    return r;
}

// E values of cell (x, p) and of its +x neighbour, with the Bloch x seam.
template <class T, bool SEAM>
__device__ __forceinline__ void load_e_pair(const Geom<T>& g, const HSys<T>& s, int x, int p, T& ex, T& ey,
                                            T& ez, T& ey1, T& ez1) {
    const size_t b = (size_t)x * g.plane + p;
    const int xn = (x == g.sx - 1) ? 0 : x + 1;
    const size_t bn = (size_t)xn * g.plane + p;
    ex = s.E[0][b];
    ey = s.E[1][b];
    ez = s.E[2][b];
    ey1 = s.E[1][bn];
    ez1 = s.E[2][bn];
    if (SEAM && (g.seam_mask & 1) && x == g.sx - 1) {
        ey1 = seam_mix(ey1, s.P[1][bn], s.sa[0], s.sb[0]);
        ez1 = seam_mix(ez1, s.P[2][bn], s.sa[0], s.sb[0]);
    }
}

// MODE: 0 = no Poynting, 1 = write oS (two-pass), 2 = fused S
template <class T, bool PAIR, bool SEAM, int MODE>
__global__ __launch_bounds__(kBlock) void h_march(HArgs<T> a) {
    const Geom<T>& g = a.g;
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= g.plane) return;
    const int y = p / g.sz;
    const int z = p - y * g.sz;
    const int ym = (y == 0) ? g.sy - 1 : y - 1;
    const int zm = (z == 0) ? g.sz - 1 : z - 1;
    const int pym = ym * g.sz + z;
    const int pzm = y * g.sz + zm;
    const int x0 = blockIdx.y * g.xchunk;
    const int x1 = min(x0 + g.xchunk, g.sx);
    const T sf = g.dt * (T)0.125;

    // register queue: E of the current plane
    T ex = a.s[0].E[0][(size_t)x0 * g.plane + p];
    T ey = a.s[0].E[1][(size_t)x0 * g.plane + p];
    T ez = a.s[0].E[2][(size_t)x0 * g.plane + p];
    T fx = (T)0, fy = (T)0, fz = (T)0;
    if (PAIR) {
        fx = a.s[1].E[0][(size_t)x0 * g.plane + p];
        fy = a.s[1].E[1][(size_t)x0 * g.plane + p];
        fz = a.s[1].E[2][(size_t)x0 * g.plane + p];
    }
    // oS of the -x neighbour (fused S)
    T oSyx_m = (T)0, oSzx_m = (T)0;
    if (MODE == 2) {
        const int xp = (x0 == 0) ? g.sx - 1 : x0 - 1;
        T qx, qy, qz, qy1, qz1;
        load_e_pair<T, SEAM>(g, a.s[0], xp, p, qx, qy, qz, qy1, qz1);
        const HCell<T> c = h_cell<T, SEAM, false>(g, a.s[0], xp, y, z, p, (size_t)xp * g.plane, qx, qy, qz,
                                                  qy1, qz1);
        const T aHy = c.hy1 + c.hy0, aHz = c.hz1 + c.hz0;
        oSyx_m = -c.aExy * aHz;
        oSzx_m = c.aExz * aHy;
    }
    for (int x = x0; x < x1; ++x) {
        const size_t b = (size_t)x * g.plane;
        const int xn = (x == g.sx - 1) ? 0 : x + 1;
        const size_t bn = (size_t)xn * g.plane + p;
        // ---- system 0 (E -> H), with Poynting
        {
            const HSys<T>& s = a.s[0];
            const T exn = s.E[0][bn], eyn = s.E[1][bn], ezn = s.E[2][bn];
            T ey1 = eyn, ez1 = ezn;
            if (SEAM && (g.seam_mask & 1) && x == g.sx - 1) {
                ey1 = seam_mix(ey1, s.P[1][bn], s.sa[0], s.sb[0]);
                ez1 = seam_mix(ez1, s.P[2][bn], s.sa[0], s.sb[0]);
            }
            const HCell<T> c = h_cell<T, SEAM, true>(g, s, x, y, z, p, b, ex, ey, ez, ey1, ez1);
            if (MODE != 0) {
                const T aHx = c.hx1 + c.hx0, aHy = c.hy1 + c.hy0, aHz = c.hz1 + c.hz0;
                const T oSxy = c.aEyx * aHz;
                const T oSxz = -c.aEzx * aHy;
                const T oSyz = c.aEzy * aHx;
                const T oSyx = -c.aExy * aHz;
                const T oSzx = c.aExz * aHy;
                const T oSzy = -c.aEyz * aHx;
                const size_t i = b + p;
                if (MODE == 1) {
                    a.oS[0][i] = oSxy; a.oS[1][i] = oSyz; a.oS[2][i] = oSzx;
                    a.oS[3][i] = oSxz; a.oS[4][i] = oSyx; a.oS[5][i] = oSzy;
                } else {
                    // -y neighbour: oSxy, oSzy; -z neighbour: oSxz, oSyz
                    T qx, qy, qz, qy1, qz1;
                    load_e_pair<T, SEAM>(g, s, x, pym, qx, qy, qz, qy1, qz1);
                    const HCell<T> cy = h_cell<T, SEAM, false>(g, s, x, ym, z, pym, b, qx, qy, qz, qy1, qz1);
                    load_e_pair<T, SEAM>(g, s, x, pzm, qx, qy, qz, qy1, qz1);
                    const HCell<T> cz = h_cell<T, SEAM, false>(g, s, x, y, zm, pzm, b, qx, qy, qz, qy1, qz1);
                    const T oSxy_ym = cy.aEyx * (cy.hz1 + cy.hz0);
                    const T oSzy_ym = -cy.aEyz * (cy.hx1 + cy.hx0);
                    const T oSxz_zm = -cz.aEzx * (cz.hy1 + cz.hy0);
                    const T oSyz_zm = cz.aEzy * (cz.hx1 + cz.hx0);
                    a.S[0][i] = (oSxy + oSxz + oSxy_ym + oSxz_zm) * sf;
                    a.S[1][i] = (oSyz + oSyx + oSyz_zm + oSyx_m) * sf;
                    a.S[2][i] = (oSzx + oSzy + oSzx_m + oSzy_ym) * sf;
                    oSyx_m = oSyx;
                    oSzx_m = oSzx;
                }
            }
            ex = exn; ey = eyn; ez = ezn;
        }
        // ---- system 1 (F -> G), no Poynting
        if (PAIR) {
            const HSys<T>& s = a.s[1];
            const T fxn = s.E[0][bn], fyn = s.E[1][bn], fzn = s.E[2][bn];
            T fy1 = fyn, fz1 = fzn;
            if (SEAM && (g.seam_mask & 1) && x == g.sx - 1) {
                fy1 = seam_mix(fy1, s.P[1][bn], s.sa[0], s.sb[0]);
                fz1 = seam_mix(fz1, s.P[2][bn], s.sa[0], s.sb[0]);
            }
            (void)h_cell<T, SEAM, true>(g, s, x, y, z, p, b, fx, fy, fz, fy1, fz1);
            fx = fxn; fy = fyn; fz = fzn;
        }
    }
}

// ---------------------------------------------------------------------------
// S from oS (two-pass reference dataflow), update_s.cl:33-36.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void s_from_os(HArgs<T> a) {
    const Geom<T>& g = a.g;
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= g.plane) return;
    const int x = blockIdx.y;
    const int y = p / g.sz;
    const int z = p - y * g.sz;
    const size_t i = (size_t)x * g.plane + p;
    const size_t imy = (size_t)x * g.plane + ((y == 0) ? p + (g.sy - 1) * g.sz : p - g.sz);
    const size_t imz = (size_t)x * g.plane + ((z == 0) ? p + g.sz - 1 : p - 1);
    const size_t imx = (size_t)((x == 0) ? g.sx - 1 : x - 1) * g.plane + p;
    const T sf = g.dt * (T)0.125;
    a.S[0][i] = (a.oS[0][i] + a.oS[3][i] + a.oS[0][imy] + a.oS[3][imz]) * sf;
    a.S[1][i] = (a.oS[1][i] + a.oS[4][i] + a.oS[1][imz] + a.oS[4][imx]) * sf;
    a.S[2][i] = (a.oS[2][i] + a.oS[5][i] + a.oS[2][imx] + a.oS[5][imy]) * sf;
}

// ---------------------------------------------------------------------------
// Field source J, update_j.cl:21-32; box given as [min, max) per axis.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void j_box(T* E, const T* Jr, const T* Ji, T c, T s, int sx, int sy,
                                                int sz, int x0, int y0, int z0, int nx, int ny, int nz) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t n = (int64_t)nx * ny * nz;
    if (t >= n) return;
    const int dz = (int)(t % nz);
    const int dy = (int)((t / nz) % ny);
    const int dx = (int)(t / ((int64_t)nz * ny));
    const size_t N = (size_t)sx * sy * sz;
    const size_t i = (size_t)(z0 + dz) + (size_t)sz * ((size_t)(y0 + dy) + (size_t)sy * (x0 + dx));
#pragma unroll
    for (int k = 0; k < 3; ++k) E[k * N + i] += c * Jr[k * N + i] + s * Ji[k * N + i];
}

template <class T>
__global__ void add_point(T* f, const T* amp) {
    if (threadIdx.x == 0) f[0] = f[0] + amp[0];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct Plan {
    fdtd_desc d;
    int sx, sy, sz, plane;
    int64_t N;
    int xchunk, nchunk, pblocks;
    int64_t psi_elems, coef_elems;
    struct F { int lo, hi, psi_off, face_elems, coef_off; } face[FDTD_MAX_FACES];
};

int64_t plane_of(const Plan& P, int axis) {
    if (axis == 0) return (int64_t)P.sy * P.sz;
    if (axis == 1) return (int64_t)P.sx * P.sz;
    return (int64_t)P.sx * P.sy;
}

template <class T>
void fill_geom(const Plan& P, const fdtd_fields* f, Geom<T>& g) {
    g.sx = P.sx; g.sy = P.sy; g.sz = P.sz; g.plane = P.plane;
    g.xchunk = P.xchunk;
    g.seam_mask = (P.d.bloch_mode == FDTD_BLOCH_PHASE) ? P.d.bloch_mask : 0;
    g.dt = (T)P.d.dt;
    g.inv_u = (T)P.d.inv_dx;
    for (int r = 0; r < 3; ++r) {
        g.inv_e[r] = nullptr;
        g.inv_h[r] = nullptr;
    }
    if (P.d.nonuniform) {
        const T* base = (const T*)f->inv_d;
        const int64_t len[3] = {P.sx, P.sy, P.sz};
        int64_t off = 0;
        for (int r = 0; r < 3; ++r) { g.inv_e[r] = base + off; off += len[r]; }
        for (int r = 0; r < 3; ++r) { g.inv_h[r] = base + off; off += len[r]; }
    }
    g.coef = (const T*)f->coef;
    memset(g.face, 0, sizeof(g.face));
    int used[3] = {0, 0, 0};
    for (int k = 0; k < P.d.n_face; ++k) {
        const int r = P.d.face[k].axis;
        FaceDev<T>& fd = g.face[r][used[r]++];
        fd.pol = P.d.face[k].polarity;
        fd.thick = P.d.face[k].thickness;
        fd.lo = P.face[k].lo;
        fd.hi = P.face[k].hi;
        fd.psi_off = (int)P.face[k].psi_off;
        fd.face_elems = (int)P.face[k].face_elems;
        fd.coef_off = P.face[k].coef_off;
    }
}

template <class T>
void comp3(void* base, int64_t N, T* out[3]) {
    for (int k = 0; k < 3; ++k) out[k] = base ? (T*)base + k * N : nullptr;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(5, "%s: launch failed: %s", what, hipGetErrorString(e));
    return 0;
}

template <class T>
int launch_e(const Plan& P, const fdtd_fields* f, hipStream_t st, const T* src_amp, int64_t src_index) {
    EArgs<T> a;
    memset(&a, 0, sizeof(a));
    fill_geom<T>(P, f, a.g);
    const int64_t N = P.N;
    T* tmp[3];
    comp3<T>(f->eps, N, tmp);
    for (int k = 0; k < 3; ++k) a.eps[k] = tmp[k];
    const bool pair = P.d.bloch_mask != 0;
    const double* re = P.d.bloch_re;
    const double* im = P.d.bloch_im;
    // system 0: E from H, partner G; seam H(-1) = (H + iG)(c - is) -> real: Hc + Gs
    comp3<T>(f->E, N, a.s[0].E);
    comp3<T>(f->H, N, tmp); for (int k = 0; k < 3; ++k) a.s[0].H[k] = tmp[k];
    comp3<T>(pair ? f->G : nullptr, N, tmp); for (int k = 0; k < 3; ++k) a.s[0].P[k] = tmp[k];
    a.s[0].psi = (T*)f->psi_e;
    for (int r = 0; r < 3; ++r) { a.s[0].sa[r] = (T)re[r]; a.s[0].sb[r] = (T)im[r]; }
    if (pair) {
        // system 1: F from G, partner H; imaginary part: Gc - Hs
        comp3<T>(f->F, N, a.s[1].E);
        comp3<T>(f->G, N, tmp); for (int k = 0; k < 3; ++k) a.s[1].H[k] = tmp[k];
        comp3<T>(f->H, N, tmp); for (int k = 0; k < 3; ++k) a.s[1].P[k] = tmp[k];
        a.s[1].psi = (T*)f->psi_f;
        for (int r = 0; r < 3; ++r) { a.s[1].sa[r] = (T)re[r]; a.s[1].sb[r] = (T)(-im[r]); }
    }
    a.src_amp = src_amp;
    if (src_amp) {
        const int64_t comp = src_index / N;
        const int64_t i = src_index - comp * N;
        a.src_comp = (int)comp;
        a.src_x = (int)(i / P.plane);
        a.src_p = (int)(i - (int64_t)a.src_x * P.plane);
    }
    dim3 grid(P.pblocks, P.nchunk);
    const bool seam = pair && P.d.bloch_mode == FDTD_BLOCH_PHASE;
    if (!pair) hipLaunchKernelGGL((e_march<T, false, false>), grid, dim3(kBlock), 0, st, a);
    else if (!seam) hipLaunchKernelGGL((e_march<T, true, false>), grid, dim3(kBlock), 0, st, a);
    else hipLaunchKernelGGL((e_march<T, true, true>), grid, dim3(kBlock), 0, st, a);
    return check_launch("fdtd_update_e");
}

template <class T, bool PAIR, bool SEAM>
void launch_h_mode(int mode, dim3 grid, hipStream_t st, const HArgs<T>& a) {
    if (mode == 0) hipLaunchKernelGGL((h_march<T, PAIR, SEAM, 0>), grid, dim3(kBlock), 0, st, a);
    else if (mode == 1) hipLaunchKernelGGL((h_march<T, PAIR, SEAM, 1>), grid, dim3(kBlock), 0, st, a);
    else hipLaunchKernelGGL((h_march<T, PAIR, SEAM, 2>), grid, dim3(kBlock), 0, st, a);
}

template <class T>
int fill_h(const Plan& P, const fdtd_fields* f, HArgs<T>& a) {
    memset(&a, 0, sizeof(a));
    fill_geom<T>(P, f, a.g);
    const int64_t N = P.N;
    const bool pair = P.d.bloch_mask != 0;
    const double* re = P.d.bloch_re;
    const double* im = P.d.bloch_im;
    T* tmp[3];
    // system 0: H from E, partner F; seam E(s) = (E + iF)(c + is) -> real: Ec - Fs
    comp3<T>(f->E, N, tmp); for (int k = 0; k < 3; ++k) a.s[0].E[k] = tmp[k];
    comp3<T>(f->H, N, tmp); for (int k = 0; k < 3; ++k) a.s[0].Hin[k] = tmp[k];
    comp3<T>(f->H_out ? f->H_out : f->H, N, a.s[0].Hout);
    comp3<T>(pair ? f->F : nullptr, N, tmp); for (int k = 0; k < 3; ++k) a.s[0].P[k] = tmp[k];
    a.s[0].psi_in = (const T*)f->psi_h;
    a.s[0].psi_out = (T*)(f->psi_h_out ? f->psi_h_out : f->psi_h);
    for (int r = 0; r < 3; ++r) { a.s[0].sa[r] = (T)re[r]; a.s[0].sb[r] = (T)(-im[r]); }
    if (pair) {
        // system 1: G from F, partner E; imaginary part: Fc + Es
        comp3<T>(f->F, N, tmp); for (int k = 0; k < 3; ++k) a.s[1].E[k] = tmp[k];
        comp3<T>(f->G, N, tmp); for (int k = 0; k < 3; ++k) a.s[1].Hin[k] = tmp[k];
        comp3<T>(f->G_out ? f->G_out : f->G, N, a.s[1].Hout);
        comp3<T>(f->E, N, tmp); for (int k = 0; k < 3; ++k) a.s[1].P[k] = tmp[k];
        a.s[1].psi_in = (const T*)f->psi_g;
        a.s[1].psi_out = (T*)(f->psi_g_out ? f->psi_g_out : f->psi_g);
        for (int r = 0; r < 3; ++r) { a.s[1].sa[r] = (T)re[r]; a.s[1].sb[r] = (T)im[r]; }
    }
    comp3<T>(f->S, N, a.S);
    for (int k = 0; k < 6; ++k) a.oS[k] = f->oS ? (T*)f->oS + k * N : nullptr;
    return 0;
}

template <class T>
int launch_h(const Plan& P, const fdtd_fields* f, hipStream_t st) {
    HArgs<T> a;
    fill_h<T>(P, f, a);
    dim3 grid(P.pblocks, P.nchunk);
    const int mode = P.d.poynting == FDTD_POYNTING_FUSED ? 2 : (P.d.poynting == FDTD_POYNTING_TWO_PASS ? 1 : 0);
    const bool pair = P.d.bloch_mask != 0;
    const bool seam = pair && P.d.bloch_mode == FDTD_BLOCH_PHASE;
    if (!pair) launch_h_mode<T, false, false>(mode, grid, st, a);
    else if (!seam) launch_h_mode<T, true, false>(mode, grid, st, a);
    else launch_h_mode<T, true, true>(mode, grid, st, a);
    return check_launch("fdtd_update_h");
}

int check_fields(const Plan& P, const fdtd_fields* f, bool for_h) {
    if (!f) return fail(2, "fields is NULL");
    if (!f->E || !f->H || !f->eps) return fail(2, "E, H and eps are required");
    if (P.d.n_face > 0 && (!f->coef || !f->psi_e || !f->psi_h))
        return fail(2, "coef, psi_e and psi_h are required with CPML faces");
    if (P.d.nonuniform && !f->inv_d) return fail(2, "inv_d is required for a nonuniform grid");
    if (P.d.bloch_mask) {
        if (!f->F || !f->G) return fail(2, "F and G are required with Bloch boundaries");
        if (P.d.n_face > 0 && (!f->psi_f || !f->psi_g))
            return fail(2, "psi_f and psi_g are required with Bloch boundaries and CPML faces");
    }
    if (for_h) {
        const void* hout = f->H_out ? f->H_out : f->H;
        if (P.d.poynting != FDTD_POYNTING_OFF && !f->S) return fail(2, "S is required with Poynting on");
        if (P.d.poynting == FDTD_POYNTING_TWO_PASS && !f->oS) return fail(2, "oS is required in two-pass mode");
        if (P.d.poynting == FDTD_POYNTING_FUSED) {
            if (hout == f->H) return fail(2, "fused Poynting needs H_out != H (ping-pong buffers)");
            if (P.d.n_face > 0 && (!f->psi_h_out || f->psi_h_out == f->psi_h))
                return fail(2, "fused Poynting needs psi_h_out != psi_h");
        }
    }
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int fdtd_abi_version(void) { return FDTD_ABI_VERSION; }

const char* fdtd_last_error(void) { return g_last_error.c_str(); }

int fdtd_plan_create(const fdtd_desc* desc, void** plan_out) {
    if (!desc || !plan_out) return fail(1, "desc and plan_out must be non-NULL");
    *plan_out = nullptr;
    const fdtd_desc& d = *desc;
    for (int r = 0; r < 3; ++r)
        if (d.shape[r] < 1) return fail(1, "shape[%d] = %lld must be >= 1", r, (long long)d.shape[r]);
    const int64_t N = d.shape[0] * d.shape[1] * d.shape[2];
    if (N > (int64_t)INT32_MAX || d.shape[1] * d.shape[2] > (int64_t)INT32_MAX)
        return fail(1, "grid of %lld cells exceeds the 2^31-1 cells per component supported", (long long)N);
    if (d.dtype != FDTD_F32 && d.dtype != FDTD_F64) return fail(1, "dtype %d unsupported (0=f32, 1=f64)", d.dtype);
    if (d.n_face < 0 || d.n_face > FDTD_MAX_FACES) return fail(1, "n_face %d out of [0, %d]", d.n_face, FDTD_MAX_FACES);
    if (!(d.dt > 0)) return fail(1, "Invalid dt: %g", d.dt);
    if (d.poynting < 0 || d.poynting > 2) return fail(1, "poynting mode %d unknown", d.poynting);
    if (d.bloch_mask < 0 || d.bloch_mask > 7) return fail(1, "bloch_mask %d out of range", d.bloch_mask);
    if (d.bloch_mode != FDTD_BLOCH_PHASE && d.bloch_mode != FDTD_BLOCH_DECOUPLED)
        return fail(1, "bloch_mode %d unknown", d.bloch_mode);
    Plan* P = new Plan();
    P->d = d;
    P->sx = (int)d.shape[0]; P->sy = (int)d.shape[1]; P->sz = (int)d.shape[2];
    P->plane = P->sy * P->sz;
    P->N = N;
    int64_t psi_off = 0, coef_off = 0;
    int per_axis[3] = {0, 0, 0};
    for (int k = 0; k < d.n_face; ++k) {
        const fdtd_face& f = d.face[k];
        if (f.axis < 0 || f.axis > 2) { delete P; return fail(1, "face %d: axis %d not in 0..2", k, f.axis); }
        if (++per_axis[f.axis] > 2) {
            delete P;
            return fail(1, "face %d: more than two CPML faces on axis %d are not supported", k, f.axis);
        }
        if (f.polarity != 0 && f.polarity != 1) { delete P; return fail(1, "face %d: polarity %d not 0/1", k, f.polarity); }
        if (f.thickness < 1) { delete P; return fail(1, "face %d: thickness %d < 1", k, f.thickness); }
        const int s = (int)d.shape[f.axis];
        Plan::F& pf = P->face[k];
        if (f.polarity == 0) { pf.lo = 0; pf.hi = f.thickness < s ? f.thickness : s; }
        else if (f.thickness <= s) { pf.lo = s - f.thickness; pf.hi = s; }
        else { pf.lo = 0; pf.hi = 0; }   // the reference's size_t test is false everywhere
        pf.face_elems = plane_of(*P, f.axis) * f.thickness;
        pf.psi_off = psi_off;
        psi_off += 2 * pf.face_elems;
        pf.coef_off = (int)coef_off;
        coef_off += 6 * f.thickness;
    }
    if (psi_off > (int64_t)INT32_MAX) { delete P; return fail(1, "psi block too large"); }
    P->psi_elems = psi_off;
    P->coef_elems = coef_off;
    P->pblocks = (P->plane + kBlock - 1) / kBlock;
    int nchunk = (kTargetBlocks + P->pblocks - 1) / P->pblocks;
    if (nchunk > P->sx) nchunk = P->sx;
    if (nchunk < 1) nchunk = 1;
    P->xchunk = (P->sx + nchunk - 1) / nchunk;
    P->nchunk = (P->sx + P->xchunk - 1) / P->xchunk;
    *plan_out = P;
    return 0;
}

int fdtd_plan_destroy(void* plan) {
    delete (Plan*)plan;
    return 0;
}

int64_t fdtd_psi_elems(const void* plan) { return plan ? ((const Plan*)plan)->psi_elems : -1; }
int64_t fdtd_coef_elems(const void* plan) { return plan ? ((const Plan*)plan)->coef_elems : -1; }

int fdtd_update_e(const void* plan, const fdtd_fields* f, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_fields(P, f, false)) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (P.d.dtype == FDTD_F32) return launch_e<float>(P, f, st, nullptr, 0);
    return launch_e<double>(P, f, st, nullptr, 0);
}

int fdtd_update_h(const void* plan, const fdtd_fields* f, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_fields(P, f, true)) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (P.d.dtype == FDTD_F32) return launch_h<float>(P, f, st);
    return launch_h<double>(P, f, st);
}

int fdtd_update_s(const void* plan, const fdtd_fields* f, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (P.d.poynting != FDTD_POYNTING_TWO_PASS) return 0;
    if (!f || !f->S || !f->oS) return fail(2, "S and oS are required");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(P.pblocks, P.sx);
    if (P.d.dtype == FDTD_F32) {
        HArgs<float> a;
        fill_h<float>(P, f, a);
        hipLaunchKernelGGL(s_from_os<float>, grid, dim3(kBlock), 0, st, a);
    } else {
        HArgs<double> a;
        fill_h<double>(P, f, a);
        hipLaunchKernelGGL(s_from_os<double>, grid, dim3(kBlock), 0, st, a);
    }
    return check_launch("fdtd_update_s");
}

int fdtd_update_j(const void* plan, const fdtd_fields* f, const void* Jr, const void* Ji, double c, double s,
                  const int64_t* box, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (!f || !f->E || !Jr || !Ji || !box) return fail(2, "E, Jr, Ji and box are required");
    // clamp the box to the grid; an empty box is a no-op (update_j.cl:23-29)
    int64_t lo[3], hi[3];
    const int64_t sz[3] = {P.sx, P.sy, P.sz};
    for (int r = 0; r < 3; ++r) {
        lo[r] = box[2 * r] < 0 ? 0 : box[2 * r];
        hi[r] = box[2 * r + 1] > sz[r] ? sz[r] : box[2 * r + 1];
        if (hi[r] <= lo[r]) return 0;
    }
    const int nx = (int)(hi[0] - lo[0]), ny = (int)(hi[1] - lo[1]), nz = (int)(hi[2] - lo[2]);
    const int64_t n = (int64_t)nx * ny * nz;
    const int blocks = (int)((n + kBlock - 1) / kBlock);
    hipStream_t st = (hipStream_t)stream;
    if (P.d.dtype == FDTD_F32)
        hipLaunchKernelGGL(j_box<float>, dim3(blocks), dim3(kBlock), 0, st, (float*)f->E, (const float*)Jr,
                           (const float*)Ji, (float)c, (float)s, P.sx, P.sy, P.sz, (int)lo[0], (int)lo[1],
                           (int)lo[2], nx, ny, nz);
    else
        hipLaunchKernelGGL(j_box<double>, dim3(blocks), dim3(kBlock), 0, st, (double*)f->E, (const double*)Jr,
                           (const double*)Ji, c, s, P.sx, P.sy, P.sz, (int)lo[0], (int)lo[1], (int)lo[2], nx,
                           ny, nz);
    return check_launch("fdtd_update_j");
}

int fdtd_add_point(const void* plan, void* field, int64_t index, const void* amp, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (!field || !amp) return fail(2, "field and amp are required");
    if (index < 0 || index >= 3 * P.N) return fail(3, "index %lld out of [0, %lld)", (long long)index, (long long)(3 * P.N));
    hipStream_t st = (hipStream_t)stream;
    if (P.d.dtype == FDTD_F32)
        hipLaunchKernelGGL(add_point<float>, dim3(1), dim3(64), 0, st, (float*)field + index, (const float*)amp);
    else
        hipLaunchKernelGGL(add_point<double>, dim3(1), dim3(64), 0, st, (double*)field + index, (const double*)amp);
    return check_launch("fdtd_add_point");
}

int fdtd_step_n(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index, const void* src_wave,
                void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (n < 0) return fail(3, "n = %d < 0", n);
    if (src_wave && (src_index < 0 || src_index >= 3 * P.N))
        return fail(3, "src_index %lld out of [0, %lld)", (long long)src_index, (long long)(3 * P.N));
    if (int rc = check_fields(P, f, true)) return rc;
    hipStream_t st = (hipStream_t)stream;
    fdtd_fields cur = *f;
    const size_t esz = P.d.dtype == FDTD_F32 ? 4 : 8;
    const bool swap = (f->H_out && f->H_out != f->H);
    for (int k = 0; k < n; ++k) {
        const void* amp = src_wave ? (const char*)src_wave + (size_t)k * esz : nullptr;
        int rc = (P.d.dtype == FDTD_F32) ? launch_e<float>(P, &cur, st, (const float*)amp, src_index)
                                         : launch_e<double>(P, &cur, st, (const double*)amp, src_index);
        if (rc) return rc;
        rc = (P.d.dtype == FDTD_F32) ? launch_h<float>(P, &cur, st) : launch_h<double>(P, &cur, st);
        if (rc) return rc;
        if (P.d.poynting == FDTD_POYNTING_TWO_PASS) {
            rc = fdtd_update_s(plan, &cur, stream);
            if (rc) return rc;
        }
        if (swap) {
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
            if (P.d.bloch_mask && cur.G_out && cur.G_out != cur.G) {
                std::swap(cur.G, cur.G_out);
                std::swap(cur.psi_g, cur.psi_g_out);
            }
        }
    }
    return 0;
}

}  // extern "C"
