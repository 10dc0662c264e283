// MI355X (gfx950) kernels and C-ABI for the FDTD time-stepping hot path.
//
// Design: a memory-bound 3-D stencil.  The reference layout is kept
// (component-major, C order, z contiguous; kernels/common.cl:34-61), so the
// (y, z) plane of one x index is contiguous in memory.  A thread owns one
// plane cell p = y*sz + z and marches along x (the slowest axis) through a
// chunk of planes:
//   * the -x (E step) / +x (H step) neighbour values stay in registers from
//     the previous plane (register queue);
//   * latency is hidden by occupancy (>= 6 waves per SIMD) rather than by
//     software pipelining, which costs more VGPRs than it gains here;
//   * 64 consecutive p of a wavefront give fully coalesced 256-byte accesses
//     for every component whatever sz is; in-plane neighbours (p - sz, p - 1,
//     p + sz, p + 1) are re-read through L1/L2;
//   * addresses are 32-bit byte offsets from per-component base pointers held
//     in SGPRs (global_load ... v_off, s_base), keeping VGPR use low;
//   * workgroup ids are remapped so that consecutive tiles run on the same
//     XCD (the hardware deals workgroups round-robin over the 8 XCDs), which
//     keeps the halo rows of neighbouring tiles in that XCD's L2.
// No MFMA: the stencil does a few flops per 4-byte word.
//
// Arithmetic follows the reference templates operation by operation
// (kernels/update_e.cl:39-107, kernels/update_h.cl:36-154,
// kernels/update_s.cl:33-36); the file is compiled with -ffp-contract=off so
// every product and sum is rounded separately, as in the CPU restatement.
//
// Poynting S is fused into the H kernel.  S at a cell needs the oS products
// of its -x, -y and -z neighbours.  The -x ones come from the previous plane
// of the march.  The in-plane ones are exchanged through LDS: a workgroup
// covers a contiguous range of plane cells preceded by sz halo cells (one
// row) whose H update it evaluates without storing it.  H and the H-side psi
// are ping-ponged in this mode, so every read of H^{n-1/2} sees the
// unmodified value whichever workgroup owns the cell.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <mutex>
#include <string>
#include <utility>

#include "../../include/fdtd.h"

namespace {

// Launch-shape tuning knobs (compile-time; the defaults are the measured best
// on MI355X for the README-scale 395x345x73 grid).
#ifndef FDTD_TARGET_BLOCKS
#define FDTD_TARGET_BLOCKS 8192      // workgroups per launch (256 CUs, several each)
#endif
#ifndef FDTD_PREFETCH
#define FDTD_PREFETCH 0              // 1: issue the loads of plane x+1 before computing plane x
                                     // (costs VGPRs/occupancy; slower on MI355X)
#endif
#ifndef FDTD_NT_E
#define FDTD_NT_E 0                  // non-temporal hint on the streamed accesses of the E kernel
#endif
#ifndef FDTD_NT_H
#define FDTD_NT_H 0                  // same for the H kernel (1 re-reads the E plane from HBM)
#endif
#ifndef FDTD_XCD
#define FDTD_XCD 1                   // XCD-contiguous workgroup order
#endif
#ifndef FDTD_HB_SMALL
#define FDTD_HB_SMALL 1024           // fused-S H workgroup for rows of <= 128 cells
#endif
#ifndef FDTD_E_WAVES
#define FDTD_E_WAVES 1               // __launch_bounds__ min waves per SIMD, E kernel
#endif
#ifndef FDTD_PREFETCH_H
#define FDTD_PREFETCH_H FDTD_PREFETCH   // same for the H kernel
#endif
#ifndef FDTD_H_TARGET
#define FDTD_H_TARGET 256            // fused-S H: workgroups resident at once (one 1024-thread workgroup
                                     // per CU); the chunking minimises rounds x chunk length
#endif
#ifndef FDTD_H_WAVES
#define FDTD_H_WAVES 6               // same, H kernel
#endif
#ifndef FDTD_HE_WAVES
#define FDTD_HE_WAVES 4              // same, fused H+E kernel
#endif
#ifndef FDTD_HE_HB_SMALL
#define FDTD_HE_HB_SMALL 1024        // fused H+E workgroup for rows of <= 128 cells
#endif
#ifndef FDTD_EH_CTAB
#define FDTD_EH_CTAB 1               // E and H kernels: CPML psi loaded with the fields, coefficients from LDS
#endif
#ifndef FDTD_HE_CTAB
#define FDTD_HE_CTAB 1               // fused H+E: CPML coefficients read from an LDS copy
#endif
#ifndef FDTD_HE_PF
#define FDTD_HE_PF 0                 // fused H+E, FP32: issue the loads of plane x+1 before computing plane x
#endif
#ifndef FDTD_HE2
#define FDTD_HE2 1                   // two-step pass available (FP32)
#endif
#ifndef FDTD_HE2_LDSCARRY
#define FDTD_HE2_LDSCARRY 1          // two-step pass: H, E, eps of the previous plane read back from LDS
#endif
#ifndef FDTD_HE2_LDSNB
#define FDTD_HE2_LDSNB 1             // two-step pass: E^n at y+1, z+1 read from LDS (written one plane ahead)
#endif
#ifndef FDTD_HE2_EARLY
#define FDTD_HE2_EARLY 0             // two-step pass: next plane's loads issued before the last stage
#endif
#ifndef FDTD_HE2_COST
#define FDTD_HE2_COST 11             // two-step pass: plane iteration cost / fused pass's, x10
#endif
#ifndef FDTD_HE2_S2
#define FDTD_HE2_S2 0                // two-step pass: also store S of its first H half-step
#endif
#ifndef FDTD_HE_TARGET
#define FDTD_HE_TARGET 256           // fused H+E: workgroups resident at once (256 CUs)
#endif
#ifndef FDTD_HE2_HB
#define FDTD_HE2_HB 1024             // two-step pass: threads per workgroup (whole rows plus three halo rows)
#endif
#ifndef FDTD_HEB_SPLIT
#define FDTD_HEB_SPLIT 0             // Bloch-pair pass: the second system's loads after the first's H update
                                     // (measured on the 64x64x256 Bloch supercell: 45.9 us per step against
                                     // 36.6 us with every load of the plane up front)
#endif
#ifndef FDTD_SEG_MIN
#define FDTD_SEG_MIN 64              // he_march / he2_march: narrowest z segment of owned cells considered
#endif
#ifndef FDTD_ROW_MAX
#define FDTD_ROW_MAX 512             // fused-S H kernel and fused H+E pass: longest rows (a 1024-thread tile
                                     // holds at least one owned row and its halo row)
#endif

constexpr int kBlock = 256;          // E kernel / H kernel without fused S
constexpr int kCoefTab = 2048;       // largest coefficient block the kernels copy to LDS
constexpr int kTargetBlocks = FDTD_TARGET_BLOCKS;

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
    int psi_off;         // element offset of psi_u inside a psi block; psi_v follows
    int face_elems;      // elements of psi_u
    int coef_off;        // element offset of p0e inside the coefficient block
};

template <class T>
struct Geom {
    int sx, sy, sz, plane;
    int xchunk, nchunk, pblocks, tile;   // launch decomposition (tile = plane cells per workgroup)
    // he_march / he2_march tiles: tyo owned rows by tzo owned z cells, tzb
    // tiles per row of the grid, tw threads per workgroup row (tw = tzo = sz:
    // whole rows, z periodic inside a row; else z segments with halo columns)
    int tyo, tzo, tzb, tw;
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
    // fused point source: E_comp[src] += *src_amp after the update
    const T* src_amp;
    int src_x, src_p, src_comp;
    // pending learned-shell correction: E += ycorr before the update on the
    // cells within ycT of a face, per component on the grid; null: none
    const float* ycorr[3];
    int ycT;
    int coef_elems;
};

template <class T>
struct HArgs {
    Geom<T> g;
    HSys<T> s[2];
    T* S[3];
    T* oS[6];            // two-pass mode: oSxy, oSyz, oSzx, oSxz, oSyx, oSzy
    int coef_elems;
};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

// 32-bit byte offset from a uniform base pointer.
template <class T>
__device__ __forceinline__ T ld(const T* base, uint32_t boff) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}
template <class T>
__device__ __forceinline__ void st(T* base, uint32_t boff, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}
// streamed accesses (each value touched once per launch); NT selects the
// non-temporal cache policy
template <bool NT, class T>
__device__ __forceinline__ T lds_(const T* base, uint32_t boff) {
    if (NT)
        return __builtin_nontemporal_load(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff));
    return ld(base, boff);
}
template <bool NT, class T>
__device__ __forceinline__ void sts_(T* base, uint32_t boff, T v) {
    if (NT)
        __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff));
    else
        st(base, boff, v);
}

// The three components of a vector field.  Split: one pointer per
// component.  Packed3: one base with the components nb bytes apart (the
// 3N layout of a field buffer), which keeps a single base pointer per field
// in scalar registers; usable while 3N values fit 32-bit byte offsets.
template <class T>
struct Packed3 {
    T* b;
    uint32_t nb;
};

template <bool NT, class T>
__device__ __forceinline__ T ldc(const T* const (&f)[3], int k, uint32_t off) { return lds_<NT>(f[k], off); }
template <bool NT, class T>
__device__ __forceinline__ T ldc(T* const (&f)[3], int k, uint32_t off) { return lds_<NT>((const T*)f[k], off); }
template <bool NT, class T>
__device__ __forceinline__ T ldc(const Packed3<const T>& f, int k, uint32_t off) {
    return lds_<NT>(f.b, off + (uint32_t)k * f.nb);
}
template <bool NT, class T>
__device__ __forceinline__ void stc(T* const (&f)[3], int k, uint32_t off, T v) { sts_<NT>(f[k], off, v); }
template <bool NT, class T>
__device__ __forceinline__ void stc(const Packed3<T>& f, int k, uint32_t off, T v) {
    sts_<NT>(f.b, off + (uint32_t)k * f.nb, v);
}

// Workgroup -> (plane tile, x chunk).  Hardware workgroup b runs on XCD
// b % 8; logical tiles are handed out so each XCD gets a contiguous range.
__device__ __forceinline__ void block_coords(int pblocks, int nchunk, int* pb, int* ck) {
    const int nb = pblocks * nchunk;
    const int b = blockIdx.x;
#if FDTD_XCD
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, j = b >> 3;
    const int lb = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + j;
#else
    const int lb = b;
    (void)nb;
#endif
    *pb = lb % pblocks;
    *ck = lb / pblocks;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also fences
// global memory, which makes every wave wait for all of its outstanding
// global stores (vmcnt(0)) at each barrier; the marching kernels exchange
// data through LDS only and never read global memory another thread of the
// launch writes, so the LDS-scope fence is all they need.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <class T>
__device__ __forceinline__ bool slab(const FaceDev<T>& f, int c, int s, int* ir) {
    if (c < f.lo || c >= f.hi) return false;
    *ir = (f.pol == 0) ? c : (s - 1) - c;
    return true;
}

// psi element index: C order over the grid shape with the face axis replaced
// by its thickness, coordinate ir along that axis.
template <int R, class T>
__device__ __forceinline__ uint32_t psi_index(const Geom<T>& g, const FaceDev<T>& f, int x, int y, int z, int p,
                                              int ir) {
    if (R == 0) return (uint32_t)ir * (uint32_t)g.plane + (uint32_t)p;
    if (R == 1) return ((uint32_t)x * (uint32_t)f.thick + (uint32_t)ir) * (uint32_t)g.sz + (uint32_t)z;
    return ((uint32_t)x * (uint32_t)g.sy + (uint32_t)y) * (uint32_t)f.thick + (uint32_t)ir;
}

// CPML of one face on the derivatives of one cell; update_e.cl:64-100 (ISE)
// and update_h.cl:80-116.  d_ab = d(field_a)/d(b).  face_step takes the
// face's coefficients at depth ir and the old psi pair (ou, ov) and stores
// the new pair at byte offsets bu, bv when write is set.
template <class T>
struct FaceCoef {
    T p0, p1, p2;
};

// A cell's psi pair per axis (u, v); only the axes whose slab holds the cell
// are set.
template <class T>
struct PsiPair {
    T u[3], v[3];
};

template <class T, bool ISE>
__device__ __forceinline__ FaceCoef<T> face_coef(const Geom<T>& g, const FaceDev<T>& f, int ir) {
    const T* c = g.coef + f.coef_off + (ISE ? 0 : 3 * f.thick);
    FaceCoef<T> k;
    k.p0 = c[ir];
    k.p1 = c[f.thick + ir];
    k.p2 = c[2 * f.thick + ir];
    return k;
}

// the same from a copy of the coefficient block (e.g. in LDS)
template <class T, bool ISE>
__device__ __forceinline__ FaceCoef<T> face_coef_at(const T* tab, const FaceDev<T>& f, int ir) {
    const T* c = tab + f.coef_off + (ISE ? 0 : 3 * f.thick);
    FaceCoef<T> k;
    k.p0 = c[ir];
    k.p1 = c[f.thick + ir];
    k.p2 = c[2 * f.thick + ir];
    return k;
}

template <int R, class T, bool ISE>
__device__ __forceinline__ void face_step(const FaceCoef<T>& k, T ou, T ov, T* psi_out, uint32_t bu,
                                          uint32_t bv, bool write, T& dxy, T& dxz, T& dyx, T& dyz, T& dzx,
                                          T& dzy, T& px, T& py, T& pz, PsiPair<T>* keep = nullptr) {
    const T p0 = k.p0, p1 = k.p1, p2 = k.p2;
    T su, sv;
    if (R == 0) {                      // r = x, (u, v) = (y, z)
        dzx *= p2; dyx *= p2;
        su = p0 * ou + p1 * dzx;
        sv = p0 * ov + p1 * dyx;
        if (ISE) { py -= su; pz += sv; } else { py += su; pz -= sv; }
    } else if (R == 1) {               // r = y, (u, v) = (x, z)
        dzy *= p2; dxy *= p2;
        su = p0 * ou + p1 * dzy;
        sv = p0 * ov + p1 * dxy;
        if (ISE) { px += su; pz -= sv; } else { px -= su; pz += sv; }
    } else {                           // r = z, (u, v) = (x, y)
        dyz *= p2; dxz *= p2;
        su = p0 * ou + p1 * dyz;
        sv = p0 * ov + p1 * dxz;
        if (ISE) { px -= su; py += sv; } else { px += su; py -= sv; }
    }
    if (write) {
        st(psi_out, bu, su);
        st(psi_out, bv, sv);
    }
    if (keep) {
        keep->u[R] = su;
        keep->v[R] = sv;
    }
}

template <int R, class T, bool ISE>
__device__ __forceinline__ void apply_face(const Geom<T>& g, const FaceDev<T>& f, int ir, uint32_t ip,
                                           const T* psi_in, T* psi_out, bool write,
                                           T& dxy, T& dxz, T& dyx, T& dyz, T& dzx, T& dzy,
                                           T& px, T& py, T& pz) {
    const FaceCoef<T> k = face_coef<T, ISE>(g, f, ir);
    const uint32_t bu = ((uint32_t)f.psi_off + ip) * (uint32_t)sizeof(T);
    const uint32_t bv = bu + (uint32_t)f.face_elems * (uint32_t)sizeof(T);
    const T ou = ld(psi_in, bu), ov = ld(psi_in, bv);
    face_step<R, T, ISE>(k, ou, ov, psi_out, bu, bv, write, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
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
template <class T, bool ISE>
__device__ __forceinline__ void cpml_cell(const Geom<T>& g, int x, int y, int z, int p, const T* psi_in,
                                          T* psi_out, bool write, T& dxy, T& dxz, T& dyx, T& dyz, T& dzx,
                                          T& dzy, T& px, T& py, T& pz) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const FaceDev<T>& f = g.face[0][k];
        int ir;
        if (f.thick && slab(f, x, g.sx, &ir))
            apply_face<0, T, ISE>(g, f, ir, psi_index<0>(g, f, x, y, z, p, ir), psi_in, psi_out, write,
                                  dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const FaceDev<T>& f = g.face[1][k];
        int ir;
        if (f.thick && slab(f, y, g.sy, &ir))
            apply_face<1, T, ISE>(g, f, ir, psi_index<1>(g, f, x, y, z, p, ir), psi_in, psi_out, write,
                                  dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const FaceDev<T>& f = g.face[2][k];
        int ir;
        if (f.thick && slab(f, z, g.sz, &ir))
            apply_face<2, T, ISE>(g, f, ir, psi_index<2>(g, f, x, y, z, p, ir), psi_in, psi_out, write,
                                  dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    }
}

// Per-thread constants of a plane cell.
struct Cell {
    int p, y, z;
    uint32_t op;          // byte offset of p
    uint32_t om_y, om_z;  // byte offsets of p - e_y, p - e_z (periodic)
    uint32_t op_y, op_z;  // byte offsets of p + e_y, p + e_z (periodic)
};

template <class T>
__device__ __forceinline__ Cell make_cell(const Geom<T>& g, int p) {
    Cell c;
    c.p = p;
    c.y = p / g.sz;
    c.z = p - c.y * g.sz;
    const int pym = (c.y == 0) ? p + (g.sy - 1) * g.sz : p - g.sz;
    const int pzm = (c.z == 0) ? p + g.sz - 1 : p - 1;
    const int pyp = (c.y == g.sy - 1) ? p - (g.sy - 1) * g.sz : p + g.sz;
    const int pzp = (c.z == g.sz - 1) ? p - (g.sz - 1) : p + 1;
    constexpr uint32_t B = sizeof(T);
    c.op = (uint32_t)p * B;
    c.om_y = (uint32_t)pym * B;
    c.om_z = (uint32_t)pzm * B;
    c.op_y = (uint32_t)pyp * B;
    c.op_z = (uint32_t)pzp * B;
    return c;
}

// A thread's place in a he_march / he2_march tile (Geom's tyo, tzo, tzb,
// tw).  Workgroup tile pb holds the tile's owned rows with YL halo rows
// below and YH above, tw threads a row, thread t in row r = t / tw, column
// cz = t % tw.  Whole rows (tzb = 1): column cz is z, z periodic inside the
// row.  Z segments (tzb > 1): the owned cells z0 .. z0+tzo-1 with ZL halo
// columns before and ZH after.  nr, ncz: the rows and columns that hold
// cells (the last tile of a column or row is cut short); in: the thread
// holds one; p: its plane cell, y and z wrapped periodically (halo cells
// may repeat cells of the tile on grids thinner than the tile).
struct TileCell {
    int p, r, cz, nr, ncz;
    bool in, seg;
};

template <int YL, int YH, int ZL, int ZH, class T>
__device__ __forceinline__ TileCell tile_cell(const Geom<T>& g, int pb, int t) {
    TileCell k;
    const int iy = pb / g.tzb;
    const int iz = pb - iy * g.tzb;
    const int y0 = iy * g.tyo;
    k.seg = g.tzb > 1;
    k.r = t / g.tw;
    k.cz = t - k.r * g.tw;
    k.nr = min(g.tyo, g.sy - y0) + YL + YH;
    int z = k.cz;
    k.ncz = g.sz;
    if (k.seg) {
        const int z0 = iz * g.tzo;
        k.ncz = min(g.tzo, g.sz - z0) + ZL + ZH;
        z = (z0 - ZL + k.cz) % g.sz;
        if (z < 0) z += g.sz;
    }
    int y = (y0 - YL + k.r) % g.sy;
    if (y < 0) y += g.sy;
    k.in = k.r < k.nr && k.cz < k.ncz;
    k.p = y * g.sz + z;
    return k;
}

// ---------------------------------------------------------------------------
// CPML with the psi pairs loaded early.  The marching kernels issue the loads
// of a cell's psi values together with its field loads at the top of a plane
// iteration (one memory round trip per plane instead of one more per slab),
// then apply the faces with cpml_apply.  A PsiHit holds, for one axis, the
// first face slot whose slab contains the cell (-1: none), the depth ir into
// it, and whether the other slot's slab contains it too (slabs of one axis
// overlap only on grids thinner than both; that second face is then applied
// with its own loads).  Same faces, same order, same arithmetic as cpml_cell.
// ---------------------------------------------------------------------------
struct PsiHit {
    int code;            // -1: none; else ir << 2 | both << 1 | slot
    __device__ __forceinline__ bool on() const { return code >= 0; }
    __device__ __forceinline__ int slot() const { return code & 1; }
    __device__ __forceinline__ bool both() const { return (code & 2) != 0; }
    __device__ __forceinline__ int ir() const { return code >> 2; }
};

template <class T>
__device__ __forceinline__ PsiHit slab_hit(const Geom<T>& g, int r, int c, int s) {
    int ir0 = 0, ir1 = 0;
    const bool h0 = g.face[r][0].thick && slab(g.face[r][0], c, s, &ir0);
    const bool h1 = g.face[r][1].thick && slab(g.face[r][1], c, s, &ir1);
    PsiHit h;
    h.code = h0 ? (ir0 << 2) | (h1 ? 2 : 0) : h1 ? (ir1 << 2) | 1 : -1;
    return h;
}

// field by field, so a per-lane slot selects between uniform values instead
// of turning into per-lane loads of the argument block
template <class T>
__device__ __forceinline__ FaceDev<T> face_of(const Geom<T>& g, int r, int slot) {
    const FaceDev<T>& a = g.face[r][0];
    const FaceDev<T>& b = g.face[r][1];
    const bool s0 = slot == 0;
    const auto pick = [s0](int u, int v) {
        return s0 ? __builtin_amdgcn_readfirstlane(u) : __builtin_amdgcn_readfirstlane(v);
    };
    FaceDev<T> f;
    f.pol = pick(a.pol, b.pol);
    f.thick = pick(a.thick, b.thick);
    f.lo = pick(a.lo, b.lo);
    f.hi = pick(a.hi, b.hi);
    f.psi_off = pick(a.psi_off, b.psi_off);
    f.face_elems = pick(a.face_elems, b.face_elems);
    f.coef_off = pick(a.coef_off, b.coef_off);
    return f;
}

template <int R, class T>
__device__ __forceinline__ uint32_t psi_byte(const Geom<T>& g, const FaceDev<T>& f, int x, const Cell& c, int ir) {
    return ((uint32_t)f.psi_off + psi_index<R>(g, f, x, c.y, c.z, c.p, ir)) * (uint32_t)sizeof(T);
}

template <int R, class T>
__device__ __forceinline__ void psi_fetch_axis(const Geom<T>& g, const T* psi_in, int x, const Cell& c,
                                               PsiHit h, PsiPair<T>& q) {
    if (h.on()) {
        const FaceDev<T> f = face_of(g, R, h.slot());
        const uint32_t bu = psi_byte<R>(g, f, x, c, h.ir());
        q.u[R] = ld(psi_in, bu);
        q.v[R] = ld(psi_in, bu + (uint32_t)f.face_elems * (uint32_t)sizeof(T));
    }
}

// hit[0..2]: the cell's PsiHit along x, y, z
template <class T>
__device__ __forceinline__ void psi_fetch(const Geom<T>& g, const T* psi_in, int x, const Cell& c,
                                          const PsiHit* hit, PsiPair<T>& q) {
    psi_fetch_axis<0>(g, psi_in, x, c, hit[0], q);
    psi_fetch_axis<1>(g, psi_in, x, c, hit[1], q);
    psi_fetch_axis<2>(g, psi_in, x, c, hit[2], q);
}

// tab: the coefficient block copied to LDS (read here, just before use), or
// null to use the coefficients k loaded beforehand
template <int R, class T, bool ISE>
__device__ __forceinline__ void cpml_axis(const Geom<T>& g, int x, const Cell& c, PsiHit h,
                                          const FaceCoef<T>& k, const T* tab, const PsiPair<T>& q,
                                          const T* psi_in, T* psi_out, bool write, T& dxy, T& dxz, T& dyx,
                                          T& dyz, T& dzx, T& dzy, T& px, T& py, T& pz,
                                          PsiPair<T>* keep = nullptr) {
    if (!h.on()) return;
    const FaceDev<T> f = face_of(g, R, h.slot());
    const uint32_t bu = psi_byte<R>(g, f, x, c, h.ir());
    const FaceCoef<T> kk = tab ? face_coef_at<T, ISE>(tab, f, h.ir()) : k;
    face_step<R, T, ISE>(kk, q.u[R], q.v[R], psi_out, bu, bu + (uint32_t)f.face_elems * (uint32_t)sizeof(T),
                         write, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz, keep);
    if (h.both()) {
        const FaceDev<T>& f1 = g.face[R][1];
        const int cr = R == 0 ? x : R == 1 ? c.y : c.z;
        const int sr = R == 0 ? g.sx : R == 1 ? g.sy : g.sz;
        int ir1 = 0;
        (void)slab(f1, cr, sr, &ir1);
        apply_face<R, T, ISE>(g, f1, ir1, psi_index<R>(g, f1, x, c.y, c.z, c.p, ir1), psi_in, psi_out, write,
                              dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    }
}

// The coefficients of the cell's slab on each axis (loaded together with
// the psi pairs, so that they share their round trip).
template <class T, bool ISE>
__device__ __forceinline__ void cpml_coefs(const Geom<T>& g, const PsiHit* hit, FaceCoef<T>* k) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        if (hit[r].on()) k[r] = face_coef<T, ISE>(g, face_of(g, r, hit[r].slot()), hit[r].ir());
    }
}

template <class T, bool ISE>
__device__ __forceinline__ void cpml_apply(const Geom<T>& g, int x, const Cell& c, const PsiHit* hit,
                                           const FaceCoef<T>* k, const T* tab, const PsiPair<T>& q,
                                           const T* psi_in, T* psi_out, bool write, T& dxy, T& dxz, T& dyx,
                                           T& dyz, T& dzx, T& dzy, T& px, T& py, T& pz,
                                           PsiPair<T>* keep = nullptr) {
    cpml_axis<0, T, ISE>(g, x, c, hit[0], k[0], tab, q, psi_in, psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px,
                         py, pz, keep);
    cpml_axis<1, T, ISE>(g, x, c, hit[1], k[1], tab, q, psi_in, psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px,
                         py, pz, keep);
    cpml_axis<2, T, ISE>(g, x, c, hit[2], k[2], tab, q, psi_in, psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px,
                         py, pz, keep);
}

// The same for an in-plane axis (y or z) with the slab data a thread needs
// worked out once before the march: the psi byte offset of the cell at
// plane x is base + x * stride (psi_index along that axis), v is vofs bytes
// further, and the coefficients of depth ir sit at cidx, cidx + thick,
// cidx + 2 * thick (E step; the H step's follow 3 * thick later).
struct PsiCol {
    PsiHit h;
    uint32_t base, stride, vofs;
    int cidx, thick;
};

template <int R, class T>
__device__ __forceinline__ PsiCol psi_col(const Geom<T>& g, const Cell& c, PsiHit h) {
    PsiCol k;
    k.h = h;
    k.base = k.stride = k.vofs = 0;
    k.cidx = k.thick = 0;
    if (h.on()) {
        const FaceDev<T> f = face_of(g, R, h.slot());
        const uint32_t B = sizeof(T);
        if (R == 1) {
            k.base = ((uint32_t)f.psi_off + (uint32_t)h.ir() * (uint32_t)g.sz + (uint32_t)c.z) * B;
            k.stride = (uint32_t)f.thick * (uint32_t)g.sz * B;
        } else {
            k.base = ((uint32_t)f.psi_off + (uint32_t)c.y * (uint32_t)f.thick + (uint32_t)h.ir()) * B;
            k.stride = (uint32_t)g.sy * (uint32_t)f.thick * B;
        }
        k.vofs = (uint32_t)f.face_elems * B;
        k.cidx = f.coef_off + h.ir();
        k.thick = f.thick;
    }
    return k;
}

template <int R, class T>
__device__ __forceinline__ void psi_fetch_col(const PsiCol& k, const T* psi_in, int x, PsiPair<T>& q) {
    if (k.h.on()) {
        const uint32_t bu = k.base + (uint32_t)x * k.stride;
        q.u[R] = ld(psi_in, bu);
        q.v[R] = ld(psi_in, bu + k.vofs);
    }
}

template <int R, class T, bool ISE>
__device__ __forceinline__ void cpml_col(const Geom<T>& g, int x, const Cell& c, const PsiCol& k, const T* tab,
                                         const PsiPair<T>& q, const T* psi_in, T* psi_out, bool write, T& dxy,
                                         T& dxz, T& dyx, T& dyz, T& dzx, T& dzy, T& px, T& py, T& pz,
                                         PsiPair<T>* keep = nullptr) {
    if (!k.h.on()) return;
    const uint32_t bu = k.base + (uint32_t)x * k.stride;
    const T* cc = tab + k.cidx + (ISE ? 0 : 3 * k.thick);
    FaceCoef<T> kk;
    kk.p0 = cc[0];
    kk.p1 = cc[k.thick];
    kk.p2 = cc[2 * k.thick];
    face_step<R, T, ISE>(kk, q.u[R], q.v[R], psi_out, bu, bu + k.vofs, write, dxy, dxz, dyx, dyz, dzx, dzy, px,
                         py, pz, keep);
    if (k.h.both()) {
        const FaceDev<T>& f1 = g.face[R][1];
        const int cr = R == 1 ? c.y : c.z;
        const int sr = R == 1 ? g.sy : g.sz;
        int ir1 = 0;
        (void)slab(f1, cr, sr, &ir1);
        apply_face<R, T, ISE>(g, f1, ir1, psi_index<R>(g, f1, x, c.y, c.z, c.p, ir1), psi_in, psi_out, write,
                              dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    }
}

// The fused pass's CPML: x from the uniform per-plane hit, y and z from the
// per-thread columns; coefficients from the table tab.  AX: the axes that
// can carry CPML faces (bit r for axis r); the code and registers of the
// other axes are left out at compile time.
template <class T, int AX = 7>
struct CpmlCols {
    PsiHit hx;
    PsiCol y, z;
};

template <class T, int AX>
__device__ __forceinline__ void psi_fetch_cols(const Geom<T>& g, const T* psi_in, int x, const Cell& c,
                                               const CpmlCols<T, AX>& k, PsiPair<T>& q) {
    if constexpr ((AX & 1) != 0) psi_fetch_axis<0>(g, psi_in, x, c, k.hx, q);
    if constexpr ((AX & 2) != 0) psi_fetch_col<1>(k.y, psi_in, x, q);
    if constexpr ((AX & 4) != 0) psi_fetch_col<2>(k.z, psi_in, x, q);
}

template <class T, bool ISE, int AX>
__device__ __forceinline__ void cpml_cols(const Geom<T>& g, int x, const Cell& c, const CpmlCols<T, AX>& k,
                                          const T* tab, const PsiPair<T>& q, const T* psi_in, T* psi_out,
                                          bool write, T& dxy, T& dxz, T& dyx, T& dyz, T& dzx, T& dzy, T& px,
                                          T& py, T& pz, PsiPair<T>* keep = nullptr) {
    const FaceCoef<T> none{};
    if constexpr ((AX & 1) != 0)
        cpml_axis<0, T, ISE>(g, x, c, k.hx, none, tab, q, psi_in, psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px,
                             py, pz, keep);
    if constexpr ((AX & 2) != 0)
        cpml_col<1, T, ISE>(g, x, c, k.y, tab, q, psi_in, psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz,
                            keep);
    if constexpr ((AX & 4) != 0)
        cpml_col<2, T, ISE>(g, x, c, k.z, tab, q, psi_in, psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz,
                            keep);
}

// CPML with prefetched psi: per-axis hits (cpml_apply) or columns (cpml_cols)
template <class T, bool ISE>
__device__ __forceinline__ void cpml_pre(const Geom<T>& g, int x, const Cell& c, const PsiHit* hit,
                                         const FaceCoef<T>* kc, const T* tab, const PsiPair<T>& q,
                                         const T* psi_in, T* psi_out, bool write, T& dxy, T& dxz, T& dyx,
                                         T& dyz, T& dzx, T& dzy, T& px, T& py, T& pz, PsiPair<T>* keep = nullptr) {
    cpml_apply<T, ISE>(g, x, c, hit, kc, tab, q, psi_in, psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz,
                       keep);
}
template <class T, bool ISE, int AX>
__device__ __forceinline__ void cpml_pre(const Geom<T>& g, int x, const Cell& c, const CpmlCols<T, AX>* cols,
                                         const FaceCoef<T>*, const T* tab, const PsiPair<T>& q,
                                         const T* psi_in, T* psi_out, bool write, T& dxy, T& dxz, T& dyx,
                                         T& dyz, T& dzx, T& dzy, T& px, T& py, T& pz, PsiPair<T>* keep = nullptr) {
    cpml_cols<T, ISE, AX>(g, x, c, *cols, tab, q, psi_in, psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz,
                          keep);
}

// ---------------------------------------------------------------------------
// E half-step, update_e.cl:20-107.
// ---------------------------------------------------------------------------

template <class T>
struct EIn {
    T hx, hy, hz, hx_ym, hz_ym, hx_zm, hy_zm, ex, ey, ez;
};

template <class T, bool SEAM>
__device__ __forceinline__ void e_fetch(const Geom<T>& g, const ESys<T>& s, const Cell& c, uint32_t ob,
                                        EIn<T>& v) {
    v.hx = ldc<FDTD_NT_E>(s.H, 0, ob + c.op);
    v.hy = ldc<FDTD_NT_E>(s.H, 1, ob + c.op);
    v.hz = ldc<FDTD_NT_E>(s.H, 2, ob + c.op);
    v.hx_ym = ldc<false>(s.H, 0, ob + c.om_y);
    v.hz_ym = ldc<false>(s.H, 2, ob + c.om_y);
    v.hx_zm = ldc<false>(s.H, 0, ob + c.om_z);
    v.hy_zm = ldc<false>(s.H, 1, ob + c.om_z);
    v.ex = ldc<FDTD_NT_E>(s.E, 0, ob + c.op);
    v.ey = ldc<FDTD_NT_E>(s.E, 1, ob + c.op);
    v.ez = ldc<FDTD_NT_E>(s.E, 2, ob + c.op);
    if constexpr (SEAM) {
        if ((g.seam_mask & 2) && c.y == 0) {
            v.hx_ym = seam_mix(v.hx_ym, ld(s.P[0], ob + c.om_y), s.sa[1], s.sb[1]);
            v.hz_ym = seam_mix(v.hz_ym, ld(s.P[2], ob + c.om_y), s.sa[1], s.sb[1]);
        }
        if ((g.seam_mask & 4) && c.z == 0) {
            v.hx_zm = seam_mix(v.hx_zm, ld(s.P[0], ob + c.om_z), s.sa[2], s.sb[2]);
            v.hy_zm = seam_mix(v.hy_zm, ld(s.P[1], ob + c.om_z), s.sa[2], s.sb[2]);
        }
    }
}

// The E update of cell (x, c) without storing E: the new E (with the source
// added) in ex, ey, ez.  psi is written in place when write is set and the
// new pairs are kept in *keep when keep is set.  QUOT: eps0..2 already hold
// the quotients dt / eps (the same IEEE division, done by the caller).
template <class T, bool PRE, class S, class K, bool QUOT = false>
__device__ __forceinline__ void e_value(const Geom<T>& g, const S& s, const Cell& c, int x, const EIn<T>& v,
                                        T hy_xm, T hz_xm, T eps0, T eps1, T eps2, T ix, T iy, T iz, int src_comp,
                                        T amp, K hit, const FaceCoef<T>* kc, const PsiPair<T>* q, const T* tab,
                                        bool write, PsiPair<T>* keep, T& ex, T& ey, T& ez) {
    T dxy = (v.hx - v.hx_ym) * iy;
    T dxz = (v.hx - v.hx_zm) * iz;
    T dyx = (v.hy - hy_xm) * ix;
    T dyz = (v.hy - v.hy_zm) * iz;
    T dzx = (v.hz - hz_xm) * ix;
    T dzy = (v.hz - v.hz_ym) * iy;
    T px = (T)0, py = (T)0, pz = (T)0;
    if (PRE)
        cpml_pre<T, true>(g, x, c, hit, kc, tab, *q, s.psi, s.psi, write, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz,
                          keep);
    else cpml_cell<T, true>(g, x, c.y, c.z, c.p, s.psi, s.psi, write, dxy, dxz, dyx, dyz, dzx, dzy, px, py, pz);
    // (locals, assigned once at the end: a conditional store through one of
    // the three references would put them in a scratch array)
    const T k0 = QUOT ? eps0 : g.dt / eps0;
    const T k1 = QUOT ? eps1 : g.dt / eps1;
    const T k2 = QUOT ? eps2 : g.dt / eps2;
    T nx = v.ex + k0 * (dzy - dyz + px);
    T ny = v.ey + k1 * (dxz - dzx + py);
    T nz = v.ez + k2 * (dyx - dxy + pz);
    if (src_comp == 0) nx = nx + amp;
    else if (src_comp == 1) ny = ny + amp;
    else if (src_comp == 2) nz = nz + amp;
    ex = nx;
    ey = ny;
    ez = nz;
}

template <class T, bool PRE = false, class S, class K = const PsiHit*>
__device__ __forceinline__ void e_update(const Geom<T>& g, const S& s, const Cell& c, int x, uint32_t oi,
                                         const EIn<T>& v, T hy_xm, T hz_xm, T eps0, T eps1, T eps2, T ix, T iy,
                                         T iz, int src_comp, T amp, K hit = nullptr,
                                         const FaceCoef<T>* kc = nullptr, const PsiPair<T>* q = nullptr,
                                         const T* tab = nullptr) {
    T ex, ey, ez;
    e_value<T, PRE>(g, s, c, x, v, hy_xm, hz_xm, eps0, eps1, eps2, ix, iy, iz, src_comp, amp, hit, kc, q, tab, true,
                    (PsiPair<T>*)nullptr, ex, ey, ez);
    stc<FDTD_NT_E>(s.E, 0, oi, ex);
    stc<FDTD_NT_E>(s.E, 1, oi, ey);
    stc<FDTD_NT_E>(s.E, 2, oi, ez);
}

// The learned shell's correction is kept per component on the grid (3 N
// values, read at the cell's own offset; only shell cells are written).  A
// cell is in the shell when x, y or z is within T of a face: in_yz is the
// per-thread part of that test for a fixed plane cell (y, z).
__device__ __forceinline__ bool shell_yz(int y, int z, int T, int sy, int sz) {
    return y < T || y >= sy - T || z < T || z >= sz - T;
}

// SHELL: add the pending learned-shell correction (EArgs::ycorr) to E
// before the update (real fields only).
// CT: the CPML coefficient block is copied to LDS and each plane's psi
// pairs are loaded with its fields (per-thread y/z slab columns).
template <class T, bool PAIR, bool SEAM, bool SHELL = false, bool CT = false>
__global__ __launch_bounds__(kBlock, FDTD_E_WAVES) void e_march(EArgs<T> a) {
    const Geom<T>& g = a.g;
    __shared__ T ctab[CT ? kCoefTab : 1];
    if (CT) {
        for (int i = threadIdx.x; i < a.coef_elems; i += kBlock) ctab[i] = g.coef[i];
        lds_barrier();
    }
    int pb, ck;
    block_coords(g.pblocks, g.nchunk, &pb, &ck);
    const int p = pb * kBlock + threadIdx.x;
    if (p >= g.plane) return;
    const Cell c = make_cell<T>(g, p);
    const uint32_t pbytes = (uint32_t)g.plane * (uint32_t)sizeof(T);
    const int x0 = ck * g.xchunk;
    const int x1 = min(x0 + g.xchunk, g.sx);
    constexpr int NS = PAIR ? 2 : 1;

    // H at the -x neighbour of the first plane (wrap; Bloch phase on x)
    const int xm = (x0 == 0) ? g.sx - 1 : x0 - 1;
    const uint32_t obm = (uint32_t)xm * pbytes;
    T hy_m[NS], hz_m[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const ESys<T>& s = a.s[k];
        hy_m[k] = ld(s.H[1], obm + c.op);
        hz_m[k] = ld(s.H[2], obm + c.op);
        if (SEAM && (g.seam_mask & 1) && x0 == 0) {
            hy_m[k] = seam_mix(hy_m[k], ld(s.P[1], obm + c.op), s.sa[0], s.sb[0]);
            hz_m[k] = seam_mix(hz_m[k], ld(s.P[2], obm + c.op), s.sa[0], s.sb[0]);
        }
    }
    const T iy = g.inv_h[1] ? g.inv_h[1][c.y] : g.inv_u;
    const T iz = g.inv_h[2] ? g.inv_h[2][c.z] : g.inv_u;
    CpmlCols<T> cols;
    if (CT) {
        cols.y = psi_col<1>(g, c, slab_hit(g, 1, c.y, g.sy));
        cols.z = psi_col<2>(g, c, slab_hit(g, 2, c.z, g.sz));
    }
    const bool in_yz = SHELL && shell_yz(c.y, c.z, a.ycT, g.sy, g.sz);

    EIn<T> cur[NS], nxt[NS];
    T e0, e1, e2, n0 = (T)0, n1 = (T)0, n2 = (T)0;
    {
        const uint32_t ob = (uint32_t)x0 * pbytes;
#pragma unroll
        for (int k = 0; k < NS; ++k) e_fetch<T, SEAM>(g, a.s[k], c, ob, cur[k]);
        e0 = lds_<FDTD_NT_E>(a.eps[0], ob + c.op);
        e1 = lds_<FDTD_NT_E>(a.eps[1], ob + c.op);
        e2 = lds_<FDTD_NT_E>(a.eps[2], ob + c.op);
    }
    for (int x = x0; x < x1; ++x) {
        const uint32_t ob = (uint32_t)x * pbytes;
        PsiPair<T> pq[NS];
        if (CT) {
            cols.hx = slab_hit(g, 0, x, g.sx);
#pragma unroll
            for (int k = 0; k < NS; ++k) psi_fetch_cols(g, (const T*)a.s[k].psi, x, c, cols, pq[k]);
        }
        // pending shell correction, loaded ahead of the plane's fields so
        // the loads overlap
        bool in_shell = false;
        float y0 = 0.f, y1 = 0.f, y2 = 0.f;
        if (SHELL) {
            in_shell = x < a.ycT || x >= g.sx - a.ycT || in_yz;
            if (in_shell) {
                const uint32_t i = (uint32_t)x * g.plane + (uint32_t)p;
                y0 = a.ycorr[0][i];
                y1 = a.ycorr[1][i];
                y2 = a.ycorr[2][i];
            }
        }
        if (!FDTD_PREFETCH && x > x0) {
#pragma unroll
            for (int k = 0; k < NS; ++k) e_fetch<T, SEAM>(g, a.s[k], c, ob, cur[k]);
            e0 = lds_<FDTD_NT_E>(a.eps[0], ob + c.op);
            e1 = lds_<FDTD_NT_E>(a.eps[1], ob + c.op);
            e2 = lds_<FDTD_NT_E>(a.eps[2], ob + c.op);
        }
        if (FDTD_PREFETCH && x + 1 < x1) {
            const uint32_t on = ob + pbytes;
#pragma unroll
            for (int k = 0; k < NS; ++k) e_fetch<T, SEAM>(g, a.s[k], c, on, nxt[k]);
            n0 = lds_<FDTD_NT_E>(a.eps[0], on + c.op);
            n1 = lds_<FDTD_NT_E>(a.eps[1], on + c.op);
            n2 = lds_<FDTD_NT_E>(a.eps[2], on + c.op);
        }
        if (SHELL && in_shell) {
            cur[0].ex = cur[0].ex + (T)y0;
            cur[0].ey = cur[0].ey + (T)y1;
            cur[0].ez = cur[0].ez + (T)y2;
        }
        const T ix = g.inv_h[0] ? g.inv_h[0][x] : g.inv_u;
        int src_comp = -1;
        T amp = (T)0;
        if (a.src_amp != nullptr && x == a.src_x && p == a.src_p) {
            src_comp = a.src_comp;
            amp = *a.src_amp;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            if constexpr (CT)
                e_update<T, true>(g, a.s[k], c, x, ob + c.op, cur[k], hy_m[k], hz_m[k], e0, e1, e2, ix, iy, iz,
                                  k == 0 ? src_comp : -1, amp, (const CpmlCols<T>*)&cols, nullptr, &pq[k],
                                  (const T*)ctab);
            else
                e_update<T>(g, a.s[k], c, x, ob + c.op, cur[k], hy_m[k], hz_m[k], e0, e1, e2, ix, iy, iz,
                            k == 0 ? src_comp : -1, amp);
            hy_m[k] = cur[k].hy;
            hz_m[k] = cur[k].hz;
            if (FDTD_PREFETCH) cur[k] = nxt[k];
        }
        if (FDTD_PREFETCH) { e0 = n0; e1 = n1; e2 = n2; }
    }
}

// ---------------------------------------------------------------------------
// H half-step, update_h.cl:22-154, with S (update_s.cl:33-36).
// ---------------------------------------------------------------------------

template <class T>
struct HCell {
    T hx0, hy0, hz0, hx1, hy1, hz1;
    T aExy, aExz, aEyx, aEyz, aEzx, aEzy;
};

// Loads of one cell's H update except E at the cell itself (register queue):
// E at x+1 (raw), E at p + e_y and p + e_z (seam-mixed), H at the cell.
template <class T>
struct HIn {
    T exn, eyn, ezn, ex_y, ez_y, ex_z, ey_z, hx, hy, hz;
};

template <class T, bool SEAM, class S>
__device__ __forceinline__ void h_fetch(const Geom<T>& g, const S& s, const Cell& c, int x, uint32_t pbytes,
                                        HIn<T>& v) {
    const uint32_t ob = (uint32_t)x * pbytes;
    const uint32_t on = (x == g.sx - 1) ? 0u : ob + pbytes;
    v.exn = ldc<FDTD_NT_H>(s.E, 0, on + c.op);
    v.eyn = ldc<FDTD_NT_H>(s.E, 1, on + c.op);
    v.ezn = ldc<FDTD_NT_H>(s.E, 2, on + c.op);
    v.ex_y = ldc<false>(s.E, 0, ob + c.op_y);
    v.ez_y = ldc<false>(s.E, 2, ob + c.op_y);
    v.ex_z = ldc<false>(s.E, 0, ob + c.op_z);
    v.ey_z = ldc<false>(s.E, 1, ob + c.op_z);
    v.hx = ldc<FDTD_NT_H>(s.Hin, 0, ob + c.op);
    v.hy = ldc<FDTD_NT_H>(s.Hin, 1, ob + c.op);
    v.hz = ldc<FDTD_NT_H>(s.Hin, 2, ob + c.op);
    if constexpr (SEAM) {
        if ((g.seam_mask & 2) && c.y == g.sy - 1) {
            v.ex_y = seam_mix(v.ex_y, ldc<false>(s.P, 0, ob + c.op_y), s.sa[1], s.sb[1]);
            v.ez_y = seam_mix(v.ez_y, ldc<false>(s.P, 2, ob + c.op_y), s.sa[1], s.sb[1]);
        }
        if ((g.seam_mask & 4) && c.z == g.sz - 1) {
            v.ex_z = seam_mix(v.ex_z, ldc<false>(s.P, 0, ob + c.op_z), s.sa[2], s.sb[2]);
            v.ey_z = seam_mix(v.ey_z, ldc<false>(s.P, 1, ob + c.op_z), s.sa[2], s.sb[2]);
        }
    }
}

// The H update of cell (x, c).  ex, ey, ez: E at the cell.  Stores H_out and
// psi_out when write is set.
template <class T, bool SEAM, bool PRE = false, class S, class K = const PsiHit*>
__device__ __forceinline__ HCell<T> h_update(const Geom<T>& g, const S& s, const Cell& c, int x,
                                             uint32_t pbytes, T ex, T ey, T ez, const HIn<T>& v, T ix, T iy,
                                             T iz, bool write, K hit = nullptr,
                                             const FaceCoef<T>* kc = nullptr, const PsiPair<T>* q = nullptr,
                                             const T* tab = nullptr, PsiPair<T>* keep = nullptr) {
    T ey1 = v.eyn, ez1 = v.ezn;
    if constexpr (SEAM) {
        if ((g.seam_mask & 1) && x == g.sx - 1) {
            ey1 = seam_mix(ey1, ldc<false>(s.P, 1, c.op), s.sa[0], s.sb[0]);
            ez1 = seam_mix(ez1, ldc<false>(s.P, 2, c.op), s.sa[0], s.sb[0]);
        }
    }
    T dxy = (v.ex_y - ex) * iy;
    T dxz = (v.ex_z - ex) * iz;
    T dyx = (ey1 - ey) * ix;
    T dyz = (v.ey_z - ey) * iz;
    T dzx = (ez1 - ez) * ix;
    T dzy = (v.ez_y - ez) * iy;
    T px = (T)0, py = (T)0, pz = (T)0;
    if (PRE)
        cpml_pre<T, false>(g, x, c, hit, kc, tab, *q, s.psi_in, s.psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px,
                           py, pz, keep);
    else
        cpml_cell<T, false>(g, x, c.y, c.z, c.p, s.psi_in, s.psi_out, write, dxy, dxz, dyx, dyz, dzx, dzy, px,
                            py, pz);
    HCell<T> r;
    r.hx0 = v.hx;
    r.hy0 = v.hy;
    r.hz0 = v.hz;
    r.hx1 = v.hx - g.dt * (dzy - dyz - px);
    r.hy1 = v.hy - g.dt * (dxz - dzx - py);
    r.hz1 = v.hz - g.dt * (dyx - dxy - pz);
    if (write) {
        const uint32_t oi = (uint32_t)x * pbytes + c.op;
        stc<FDTD_NT_H>(s.Hout, 0, oi, r.hx1);
        stc<FDTD_NT_H>(s.Hout, 1, oi, r.hy1);
        stc<FDTD_NT_H>(s.Hout, 2, oi, r.hz1);
    }
    r.aExy = v.ex_y + ex; r.aExz = v.ex_z + ex;
    r.aEyx = ey1 + ey; r.aEyz = v.ey_z + ey;
    r.aEzx = ez1 + ez; r.aEzy = v.ez_y + ez;
    return r;
}

// The H update of an arbitrary cell, evaluated without storing anything.
template <class T, bool SEAM>
__device__ __forceinline__ HCell<T> h_eval(const Geom<T>& g, const HSys<T>& s, int x, int p, uint32_t pbytes) {
    const Cell c = make_cell<T>(g, p);
    HIn<T> v;
    h_fetch<T, SEAM>(g, s, c, x, pbytes, v);
    const uint32_t oi = (uint32_t)x * pbytes + c.op;
    const T ex = ld(s.E[0], oi), ey = ld(s.E[1], oi), ez = ld(s.E[2], oi);
    const T ix = g.inv_e[0] ? g.inv_e[0][x] : g.inv_u;
    const T iy = g.inv_e[1] ? g.inv_e[1][c.y] : g.inv_u;
    const T iz = g.inv_e[2] ? g.inv_e[2][c.z] : g.inv_u;
    return h_update<T, SEAM>(g, s, c, x, pbytes, ex, ey, ez, v, ix, iy, iz, false);
}

// MODE: 0 = no Poynting, 1 = write oS (two-pass), 2 = fused S with an LDS
// exchange, 3 = fused S with the in-plane neighbours re-evaluated per cell
// (any sz).
//
// MODE 2: a workgroup of HB threads owns g.tile plane cells, a whole number
// of rows (tile = k*sz), plus one halo row in front: the row before its first
// row, taken cyclically, so workgroup 0's halo is row sy-1.  Every -y and -z
// neighbour of an owned cell (including the periodic wraps y=0 -> sy-1 and
// z=0 -> sz-1) is then evaluated by a thread of the same workgroup.
//
// Fused modes start the march one plane early: that first iteration only
// evaluates the oS products of plane x0-1 (the -x neighbours of plane x0).
// CT (modes 0-2): the CPML coefficient block is copied to LDS and each
// plane's psi pairs are loaded with its fields (per-thread y/z columns).
template <class T, bool PAIR, bool SEAM, int MODE, int HB, bool CT = false>
__global__ __launch_bounds__(HB, HB >= 1024 ? 4 : FDTD_H_WAVES) void h_march(HArgs<T> a) {
    constexpr bool LDSX = (MODE == 2);
    constexpr bool FUSED = (MODE == 2 || MODE == 3);
    __shared__ T lds[LDSX ? 2 : 1][4][LDSX ? HB : 1];
    __shared__ T ctab[CT ? kCoefTab : 1];
    const Geom<T>& g = a.g;
    if (CT) {
        for (int i = threadIdx.x; i < a.coef_elems; i += HB) ctab[i] = g.coef[i];
        lds_barrier();
    }
    int pb, ck;
    block_coords(g.pblocks, g.nchunk, &pb, &ck);
    const int t = threadIdx.x;
    const int halo = LDSX ? g.sz : 0;
    const int praw = pb * g.tile - halo + t;
    const bool valid = (t < g.tile + halo) && praw < g.plane;
    if (!LDSX && !valid) return;   // no workgroup barrier outside the LDS mode
    const int p = praw < 0 ? praw + g.plane : praw;
    const bool own = valid && t >= halo;
    const Cell c = make_cell<T>(g, valid ? p : 0);
    const uint32_t pbytes = (uint32_t)g.plane * (uint32_t)sizeof(T);
    const int x0 = ck * g.xchunk;
    const int x1 = min(x0 + g.xchunk, g.sx);
    const T sf = g.dt * (T)0.125;
    const T iy = g.inv_e[1] ? g.inv_e[1][c.y] : g.inv_u;
    const T iz = g.inv_e[2] ? g.inv_e[2][c.z] : g.inv_u;
    constexpr int NS = PAIR ? 2 : 1;
    const int xs = FUSED ? (x0 == 0 ? g.sx - 1 : x0 - 1) : x0;
    const int n_it = (x1 - x0) + (FUSED ? 1 : 0);
    CpmlCols<T> cols;
    if (CT) {
        cols.y = psi_col<1>(g, c, slab_hit(g, 1, c.y, g.sy));
        cols.z = psi_col<2>(g, c, slab_hit(g, 2, c.z, g.sz));
    }

    // register queue: E (and F) at the current plane
    T qx[NS], qy[NS], qz[NS];
    HIn<T> cur[NS], nxt[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const uint32_t ob = (uint32_t)xs * pbytes;
        qx[k] = ld(a.s[k].E[0], ob + c.op);
        qy[k] = ld(a.s[k].E[1], ob + c.op);
        qz[k] = ld(a.s[k].E[2], ob + c.op);
        h_fetch<T, SEAM>(g, a.s[k], c, xs, pbytes, cur[k]);
    }
    T oSyx_m = (T)0, oSzx_m = (T)0;
    int x = xs;
    for (int it = 0; it < n_it; ++it) {
        const int xn = (x == g.sx - 1) ? 0 : x + 1;
        if (!FDTD_PREFETCH_H && it > 0) {
#pragma unroll
            for (int k = 0; k < NS; ++k) h_fetch<T, SEAM>(g, a.s[k], c, x, pbytes, cur[k]);
        }
        if (FDTD_PREFETCH_H && it + 1 < n_it) {
#pragma unroll
            for (int k = 0; k < NS; ++k) h_fetch<T, SEAM>(g, a.s[k], c, xn, pbytes, nxt[k]);
        }
        PsiPair<T> pq[NS];
        if (CT) {
            cols.hx = slab_hit(g, 0, x, g.sx);
            if (valid) {
#pragma unroll
                for (int k = 0; k < NS; ++k) psi_fetch_cols(g, a.s[k].psi_in, x, c, cols, pq[k]);
            }
        }
        const bool dry = FUSED && it == 0;   // uniform
        const T ix = g.inv_e[0] ? g.inv_e[0][x] : g.inv_u;
        const uint32_t oi = (uint32_t)x * pbytes + c.op;
        // ---- system 0 (E -> H), with Poynting
        T oSxy = (T)0, oSxz = (T)0, oSyz = (T)0, oSyx = (T)0, oSzx = (T)0, oSzy = (T)0;
        if (valid) {
            HCell<T> h;
            if constexpr (CT)
                h = h_update<T, SEAM, true>(g, a.s[0], c, x, pbytes, qx[0], qy[0], qz[0], cur[0], ix, iy, iz,
                                            own && !dry, (const CpmlCols<T>*)&cols, nullptr, &pq[0],
                                            (const T*)ctab);
            else
                h = h_update<T, SEAM>(g, a.s[0], c, x, pbytes, qx[0], qy[0], qz[0], cur[0], ix, iy, iz,
                                      own && !dry);
            if (MODE != 0) {
                const T aHx = h.hx1 + h.hx0, aHy = h.hy1 + h.hy0, aHz = h.hz1 + h.hz0;
                oSxy = h.aEyx * aHz;
                oSxz = -h.aEzx * aHy;
                oSyz = h.aEzy * aHx;
                oSyx = -h.aExy * aHz;
                oSzx = h.aExz * aHy;
                oSzy = -h.aEyz * aHx;
            }
            if (MODE == 1) {
                st(a.oS[0], oi, oSxy); st(a.oS[1], oi, oSyz); st(a.oS[2], oi, oSzx);
                st(a.oS[3], oi, oSxz); st(a.oS[4], oi, oSyx); st(a.oS[5], oi, oSzy);
            }
        }
        if (FUSED && !dry) {
            // oSxy, oSzy of the -y neighbour; oSxz, oSyz of the -z neighbour
            T nxy = (T)0, nzy = (T)0, nxz = (T)0, nyz = (T)0;
            if (LDSX) {
                const int b = it & 1;
                lds[b][0][t] = oSxy;
                lds[b][1][t] = oSzy;
                lds[b][2][t] = oSxz;
                lds[b][3][t] = oSyz;
                lds_barrier();
                if (own) {
                    nxy = lds[b][0][t - g.sz];
                    nzy = lds[b][1][t - g.sz];
                    const int tz = (c.z > 0) ? t - 1 : t + g.sz - 1;
                    nxz = lds[b][2][tz];
                    nyz = lds[b][3][tz];
                }
            } else if (own) {
                const int pym = (c.y == 0) ? p + (g.sy - 1) * g.sz : p - g.sz;
                const HCell<T> hy = h_eval<T, SEAM>(g, a.s[0], x, pym, pbytes);
                nxy = hy.aEyx * (hy.hz1 + hy.hz0);
                nzy = -hy.aEyz * (hy.hx1 + hy.hx0);
                const int pzm = (c.z == 0) ? p + g.sz - 1 : p - 1;
                const HCell<T> hz = h_eval<T, SEAM>(g, a.s[0], x, pzm, pbytes);
                nxz = -hz.aEzx * (hz.hy1 + hz.hy0);
                nyz = hz.aEzy * (hz.hx1 + hz.hx0);
            }
            if (own) {
                sts_<FDTD_NT_H>(a.S[0], oi, (oSxy + oSxz + nxy + nxz) * sf);
                sts_<FDTD_NT_H>(a.S[1], oi, (oSyz + oSyx + nyz + oSyx_m) * sf);
                sts_<FDTD_NT_H>(a.S[2], oi, (oSzx + oSzy + oSzx_m + nzy) * sf);
            }
        }
        oSyx_m = oSyx;
        oSzx_m = oSzx;
        // ---- system 1 (F -> G), no Poynting
        if (PAIR && own && !dry) {
            if constexpr (CT)
                (void)h_update<T, SEAM, true>(g, a.s[1], c, x, pbytes, qx[1], qy[1], qz[1], cur[1], ix, iy, iz,
                                              true, (const CpmlCols<T>*)&cols, nullptr, &pq[NS - 1],
                                              (const T*)ctab);
            else
                (void)h_update<T, SEAM>(g, a.s[1], c, x, pbytes, qx[1], qy[1], qz[1], cur[1], ix, iy, iz, true);
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            qx[k] = cur[k].exn;
            qy[k] = cur[k].eyn;
            qz[k] = cur[k].ezn;
            if (FDTD_PREFETCH_H) cur[k] = nxt[k];
        }
        x = xn;
    }
}

// ---------------------------------------------------------------------------
// H half-step with S, then the next E half-step, in one march (he_march).
//
// One pass over the grid instead of two: E^n, H^{n-1/2} and eps are read
// once, H^{n+1/2}, S and E^{n+1} written once (72 bytes per FP32 cell instead
// of 96).  The workgroup layout is MODE 2's: whole rows plus one halo row,
// and a dry plane x0-1 in front of the chunk.  At plane x every thread
// evaluates H^{n+1/2}(x) and the oS products; they go through LDS, and the
// owned cells then finish S(x) and do the E update of plane x, which needs
// H^{n+1/2} at x (own), x-1 (previous iteration, registers), y-1 (the row
// before, or the halo row) and z-1 (same row) -- the neighbours whose oS
// products S needs too.  E is read from E and written to E_out, H from H and
// written to H_out, so no workgroup ever reads a value another one has
// already overwritten; psi_E is touched by its own cell only (in place).
// Per cell the arithmetic is h_update + S (as in h_march MODE 2) followed by
// e_update (as in e_march), so the results are bit-identical to the two
// kernels run one after the other.
//
// What keeps the march short per plane (measured on MI355X):
//   * every load of a plane is issued before any of them is used: the field
//     loads, eps, and the psi pairs of both half-steps (one memory round trip
//     per plane; loading a slab face's psi inside its branch costs one more
//     round trip per face);
//   * y/z CPML slab offsets per thread are worked out once (psi_col), the
//     coefficient block is read from an LDS copy, and each field is one
//     base pointer plus 32-bit offsets (Packed3): few scalar registers, no
//     spills;
//   * the barrier orders LDS only (lds_barrier), so stores stay in flight;
//   * one 1024-thread workgroup per CU, chunks chosen by make_round_launch.
// ---------------------------------------------------------------------------
// The fused pass's systems: packed fields, no Bloch partner.
template <class T>
struct HSysP {
    Packed3<const T> E, Hin;
    Packed3<T> Hout;
    const T* psi_in;
    T* psi_out;
};

template <class T>
struct ESysP {
    Packed3<T> E;            // written: E_out
    T* psi;
};

template <class T>
struct HEArgs {
    Geom<T> g;
    HSysP<T> hs;
    ESysP<T> es;
    Packed3<T> S;
    Packed3<const T> eps;
    const T* src_amp;        // E_out[src] += *src_amp after the E update; null: none
    int src_x, src_p, src_comp;
    int coef_elems;
    // learned shell (SHELL): the correction added to E^n before its E update
    // on the cells within ycT of a face, per component on the grid
    Packed3<const float> ycorr;
    int ycT;
};

// NU: nonuniform grid (per-index 1/dx arrays); uniform grids keep those
// pointers out of the kernel's scalar registers.
// CT: the coefficient block (at most kCoefTab values) is copied to LDS
// once and read there.

// SHELL: add the pending learned-shell correction to E^n before the E
// update (the H update and S use E^n as it is).
// POY: fused Poynting (S from the oS products of H^{n+1/2}); without it the
// pass skips the products and S (do_poynting off).
template <class T, int HB, bool NU, bool CT, bool SHELL, bool POY = true,
          bool PF = (FDTD_HE_PF != 0 && sizeof(T) == 4)>
__global__ __launch_bounds__(HB, FDTD_HE_WAVES) void he_march(HEArgs<T> a) {
    __shared__ T lds[2][7][HB];
    __shared__ T ctab[CT ? kCoefTab : 1];
    const Geom<T>& g = a.g;
    const HSysP<T>& hs = a.hs;
    const ESysP<T>& es = a.es;
    int pb, ck;
    block_coords(g.pblocks, g.nchunk, &pb, &ck);
    const int t = threadIdx.x;
    // one halo row below the owned rows (and, with z segments, one halo
    // column before the owned cells): the -y / -z neighbours of the E update
    bool valid, own;
    int p;
    if (g.tzb == 1) {
        // whole rows: the tile and its halo row are one contiguous range of
        // the plane (the cells tile_cell gives).  Measured on the README
        // grid: with tile_cell here the compiler makes the next plane's
        // loads wait for this plane's stores (s_waitcnt vmcnt before them)
        // and the pass takes 227 instead of 197 us.
        const int halo = g.sz;
        const int praw = pb * g.tile - halo + t;
        valid = (t < g.tile + halo) && praw < g.plane;
        p = praw < 0 ? praw + g.plane : praw;
        own = valid && t >= halo;
    } else {
        const TileCell tc = tile_cell<1, 0, 1, 0>(g, pb, t);
        valid = tc.in;
        p = tc.p;
        own = valid && tc.r >= 1 && tc.cz >= 1;
    }
    const Cell c = make_cell<T>(g, valid ? p : 0);
    const uint32_t pbytes = (uint32_t)g.plane * (uint32_t)sizeof(T);
    const int x0 = ck * g.xchunk;
    const int x1 = min(x0 + g.xchunk, g.sx);
    const T sf = g.dt * (T)0.125;
    const T iy = NU ? g.inv_e[1][c.y] : g.inv_u;     // H step
    const T iz = NU ? g.inv_e[2][c.z] : g.inv_u;
    const T jy = NU ? g.inv_h[1][c.y] : g.inv_u;     // E step
    const T jz = NU ? g.inv_h[2][c.z] : g.inv_u;
    const int xs = (x0 == 0) ? g.sx - 1 : x0 - 1;
    const int n_it = (x1 - x0) + 1;
    const int src_comp = (own && a.src_amp != nullptr && p == a.src_p) ? a.src_comp : -1;
    if (CT) {
        for (int i = t; i < a.coef_elems; i += HB) ctab[i] = g.coef[i];
        lds_barrier();
    }
    const T* tab = CT ? ctab : nullptr;
    PsiHit hit[3];
    hit[1] = slab_hit(g, 1, c.y, g.sy);
    hit[2] = slab_hit(g, 2, c.z, g.sz);
    // with the table in LDS: the y and z slab data per thread, worked out once
    CpmlCols<T> cols;
    if (CT) {
        cols.y = psi_col<1>(g, c, hit[1]);
        cols.z = psi_col<2>(g, c, hit[2]);
    }

    T qx, qy, qz;                 // E^n at the current plane
    {
        const uint32_t ob = (uint32_t)xs * pbytes;
        qx = ldc<false>(hs.E, 0, ob + c.op);
        qy = ldc<false>(hs.E, 1, ob + c.op);
        qz = ldc<false>(hs.E, 2, ob + c.op);
    }
    // Everything a plane iteration loads from memory: the H-step inputs,
    // the psi pairs of both half-steps and eps.  With PF the loads of plane
    // x+1 are issued before plane x is computed.
    const bool in_yz = SHELL && shell_yz(c.y, c.z, a.ycT, g.sy, g.sz);
    struct PlaneIn {
        HIn<T> h;
        PsiPair<T> qh, qe;    // (left unset where the cell is in no slab: only read where it is)
        T e0, e1, e2;
        PsiHit hx;
        float y0, y1, y2;     // pending shell correction (SHELL)
        bool ins;
    };
    const auto fetch = [&](int xf, bool dryf, PlaneIn& in) {
        const uint32_t of = (uint32_t)xf * pbytes + c.op;
        PsiHit ht[3] = {slab_hit(g, 0, xf, g.sx), hit[1], hit[2]};
        in.hx = ht[0];
        in.e0 = in.e1 = in.e2 = (T)1;
        CpmlCols<T> kc = cols;
        kc.hx = ht[0];
        if (valid) {
            if (CT) psi_fetch_cols(g, hs.psi_in, xf, c, kc, in.qh);
            else psi_fetch(g, hs.psi_in, xf, c, ht, in.qh);
        }
        if (own && !dryf) {
            if (CT) psi_fetch_cols(g, (const T*)es.psi, xf, c, kc, in.qe);
            else psi_fetch(g, (const T*)es.psi, xf, c, ht, in.qe);
            in.e0 = ldc<FDTD_NT_E>(a.eps, 0, of);
            in.e1 = ldc<FDTD_NT_E>(a.eps, 1, of);
            in.e2 = ldc<FDTD_NT_E>(a.eps, 2, of);
        }
        in.ins = false;
        if (SHELL && own && !dryf) {
            in.ins = xf < a.ycT || xf >= g.sx - a.ycT || in_yz;
            if (in.ins) {
                in.y0 = ldc<false>(a.ycorr, 0, of);
                in.y1 = ldc<false>(a.ycorr, 1, of);
                in.y2 = ldc<false>(a.ycorr, 2, of);
            }
        }
        h_fetch<T, false>(g, hs, c, xf, pbytes, in.h);
    };
    PlaneIn cur;
    if (PF) fetch(xs, true, cur);
    T oSyx_m = (T)0, oSzx_m = (T)0;   // oS products of the -x neighbour
    T hy_m = (T)0, hz_m = (T)0;       // H^{n+1/2} of the -x neighbour
    int x = xs;
    for (int it = 0; it < n_it; ++it) {
        const int xn = (x == g.sx - 1) ? 0 : x + 1;
        const bool dry = it == 0;   // uniform
        const uint32_t oi = (uint32_t)x * pbytes + c.op;
        PlaneIn nxt;
        if (PF) {
            if (it + 1 < n_it) fetch(xn, false, nxt);
        } else {
            fetch(x, dry, cur);
        }
        hit[0] = cur.hx;
        cols.hx = cur.hx;
        FaceCoef<T> kh[3], ke[3];
        const T jx = NU ? g.inv_h[0][x] : g.inv_u;
        const bool src = src_comp >= 0 && x == a.src_x;
        const T amp = src ? *a.src_amp : (T)0;
        if (!CT) {
            cpml_coefs<T, false>(g, hit, kh);
            cpml_coefs<T, true>(g, hit, ke);
        }
        const auto kcpml = [&]() {
            if constexpr (CT) return (const CpmlCols<T>*)&cols;
            else return (const PsiHit*)hit;
        }();
        const T ix = NU ? g.inv_e[0][x] : g.inv_u;
        T oSxy = (T)0, oSxz = (T)0, oSyz = (T)0, oSyx = (T)0, oSzx = (T)0, oSzy = (T)0;
        T hx1 = (T)0, hy1 = (T)0, hz1 = (T)0;
        if (valid) {
            const HCell<T> h = h_update<T, false, true>(g, hs, c, x, pbytes, qx, qy, qz, cur.h, ix, iy, iz,
                                                        own && !dry, kcpml, kh, &cur.qh, tab);
            if constexpr (POY) {
                const T aHx = h.hx1 + h.hx0, aHy = h.hy1 + h.hy0, aHz = h.hz1 + h.hz0;
                oSxy = h.aEyx * aHz;
                oSxz = -h.aEzx * aHy;
                oSyz = h.aEzy * aHx;
                oSyx = -h.aExy * aHz;
                oSzx = h.aExz * aHy;
                oSzy = -h.aEyz * aHx;
            }
            hx1 = h.hx1; hy1 = h.hy1; hz1 = h.hz1;
        }
        const int b = it & 1;
        if constexpr (POY) {
            lds[b][0][t] = oSxy;
            lds[b][1][t] = oSzy;
            lds[b][2][t] = oSxz;
            lds[b][3][t] = oSyz;
        }
        lds[b][4][t] = hx1;
        lds[b][5][t] = hy1;
        lds[b][6][t] = hz1;
        lds_barrier();
        if (own && !dry) {
            // LDS slots of the -y and -z neighbours: owned cells of a z
            // segment have their -z neighbour in the column before, a whole
            // row wraps inside the row
            const int ty = t - g.tw;
            const int tz = (g.tzb > 1 || c.z > 0) ? t - 1 : t + g.tw - 1;
            if constexpr (POY) {
                const T nxy = lds[b][0][ty], nzy = lds[b][1][ty];
                const T nxz = lds[b][2][tz], nyz = lds[b][3][tz];
                stc<FDTD_NT_H>(a.S, 0, oi, (oSxy + oSxz + nxy + nxz) * sf);
                stc<FDTD_NT_H>(a.S, 1, oi, (oSyz + oSyx + nyz + oSyx_m) * sf);
                stc<FDTD_NT_H>(a.S, 2, oi, (oSzx + oSzy + oSzx_m + nzy) * sf);
            }
            EIn<T> v;
            v.hx = hx1; v.hy = hy1; v.hz = hz1;
            v.hx_ym = lds[b][4][ty];
            v.hz_ym = lds[b][6][ty];
            v.hx_zm = lds[b][4][tz];
            v.hy_zm = lds[b][5][tz];
            v.ex = qx; v.ey = qy; v.ez = qz;
            if (SHELL && cur.ins) {
                v.ex = v.ex + (T)cur.y0;
                v.ey = v.ey + (T)cur.y1;
                v.ez = v.ez + (T)cur.y2;
            }
            e_update<T, true>(g, es, c, x, oi, v, hy_m, hz_m, cur.e0, cur.e1, cur.e2, jx, jy, jz,
                              src ? src_comp : -1, amp, kcpml, ke, &cur.qe, tab);
        }
        oSyx_m = oSyx;
        oSzx_m = oSzx;
        hy_m = hy1;
        hz_m = hz1;
        qx = cur.h.exn;
        qy = cur.h.eyn;
        qz = cur.h.ezn;
        if (PF) cur = nxt;
        x = xn;
    }
}

// ---------------------------------------------------------------------------
// The fused pass for a Bloch pair (heb_march): he_march's H+S then E pass for
// the two coupled systems at once, E/H (with S) and their Bloch partners F/G
// (simulation.py:187-196, 214-224, 252-264 launch four kernels per step for
// them).  Per plane x, every thread of the tile evaluates H^{n+1/2} and
// G^{n+1/2} (seam-mixed E/F neighbours across x, y and z seams as
// h_march<PAIR, SEAM>); the oS products of the E/H system and both new H
// and G go through LDS, then the owned cells finish S and update E and F,
// mixing H and G of the -y / -z neighbours (LDS) and of the -x neighbour
// (registers) across a seam as e_march<PAIR, SEAM> does.  Every buffer is
// read from a copy the pass does not write: E -> E_out, F -> F_out,
// H -> H_out, G -> G_out, psi_h -> psi_h_out, psi_g -> psi_g_out (psi_e and
// psi_f, touched only by their own cell, in place).  Per cell the
// arithmetic is that of h_update + S and e_update for each system, in the
// same order, so the results are bit-identical to the paired h_march and
// e_march.  FP32 with the coefficient block in LDS.
// ---------------------------------------------------------------------------
template <class T>
struct HSysB {
    Packed3<const T> E, Hin, P;     // P: the partner of E across a seam
    Packed3<T> Hout;
    const T* psi_in;
    T* psi_out;
    T sa[3], sb[3];                 // seam neighbour = sa*own + sb*partner
};

template <class T>
struct HEBArgs {
    Geom<T> g;
    HSysB<T> hs[2];                 // 0: H from E (partner F), 1: G from F (partner E)
    ESysP<T> es[2];                 // 0: E_out, psi_e; 1: F_out, psi_f
    T esa[2][3], esb[2][3];         // E-step seams: 0: H with partner G, 1: G with partner H
    Packed3<T> S;
    Packed3<const T> eps;
    const T* src_amp;               // E_out[src] += *src_amp after the E update; null: none
    int src_x, src_p, src_comp;
    int coef_elems;
};

template <class T, bool SEAM, int AX>
__global__ __launch_bounds__(1024, FDTD_HE_WAVES) void heb_march(HEBArgs<T> a) {
    constexpr int HB = 1024;
    __shared__ T lds[2][10][HB];
    __shared__ T ctab[kCoefTab];
    const Geom<T>& g = a.g;
    int pb, ck;
    block_coords(g.pblocks, g.nchunk, &pb, &ck);
    const int t = threadIdx.x;
    const int halo = g.sz;
    const int praw = pb * g.tile - halo + t;
    const bool valid = (t < g.tile + halo) && praw < g.plane;
    const int p = praw < 0 ? praw + g.plane : praw;
    const bool own = valid && t >= halo;
    const Cell c = make_cell<T>(g, valid ? p : 0);
    const uint32_t pbytes = (uint32_t)g.plane * (uint32_t)sizeof(T);
    const int x0 = ck * g.xchunk;
    const int x1 = min(x0 + g.xchunk, g.sx);
    const T sf = g.dt * (T)0.125;
    const T iu = g.inv_u;
    const int xs = (x0 == 0) ? g.sx - 1 : x0 - 1;
    const int n_it = (x1 - x0) + 1;
    const int src_comp = (own && a.src_amp != nullptr && p == a.src_p) ? a.src_comp : -1;
    for (int i = t; i < a.coef_elems; i += HB) ctab[i] = g.coef[i];
    lds_barrier();
    const T* tab = ctab;
    CpmlCols<T, AX> cols{};
    cols.hx.code = cols.y.h.code = cols.z.h.code = -1;
    if constexpr ((AX & 2) != 0) cols.y = psi_col<1>(g, c, slab_hit(g, 1, c.y, g.sy));
    if constexpr ((AX & 4) != 0) cols.z = psi_col<2>(g, c, slab_hit(g, 2, c.z, g.sz));
    const int ty = t - g.sz;
    const int tz = (c.z > 0) ? t - 1 : t + g.sz - 1;

    T q[2][3];                      // E^n and F^n at the current plane
    {
        const uint32_t ob = (uint32_t)xs * pbytes + c.op;
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int r = 0; r < 3; ++r) q[k][r] = ldc<false>(a.hs[k].E, r, ob);
    }
    T oSyx_m = (T)0, oSzx_m = (T)0;           // oS products of the -x neighbour
    T hym[2] = {(T)0, (T)0}, hzm[2] = {(T)0, (T)0};   // H^{n+1/2}, G^{n+1/2} of the -x neighbour
    int x = xs;
    for (int it = 0; it < n_it; ++it) {
        const int xn = (x == g.sx - 1) ? 0 : x + 1;
        const bool dry = it == 0;   // uniform
        const uint32_t oi = (uint32_t)x * pbytes + c.op;
        // ---- loads of the plane: both H-step inputs, the four psi pairs, eps
        if constexpr ((AX & 1) != 0) cols.hx = slab_hit(g, 0, x, g.sx);
        HIn<T> in[2];
        PsiPair<T> qh, qe[2];
        T e0 = (T)1, e1 = (T)1, e2 = (T)1;
        // ---- H^{n+1/2} and G^{n+1/2} on every row of the tile; oS of E/H.
        // FDTD_HEB_SPLIT: the second system's loads (and the E-step loads)
        // are issued after the first system's H update, which keeps fewer
        // values live at once.
        T h1[2][3] = {{(T)0, (T)0, (T)0}, {(T)0, (T)0, (T)0}};
        T oSxy = (T)0, oSxz = (T)0, oSyz = (T)0, oSyx = (T)0, oSzx = (T)0, oSzy = (T)0;
        const auto fetch_e = [&]() {
            if (own && !dry) {
                psi_fetch_cols(g, (const T*)a.es[0].psi, x, c, cols, qe[0]);
                psi_fetch_cols(g, (const T*)a.es[1].psi, x, c, cols, qe[1]);
                e0 = ldc<FDTD_NT_E>(a.eps, 0, oi);
                e1 = ldc<FDTD_NT_E>(a.eps, 1, oi);
                e2 = ldc<FDTD_NT_E>(a.eps, 2, oi);
            }
        };
        if (!FDTD_HEB_SPLIT) {
            h_fetch<T, SEAM>(g, a.hs[1], c, x, pbytes, in[1]);
            fetch_e();
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (k == 0 || FDTD_HEB_SPLIT) {
                if (valid) psi_fetch_cols(g, a.hs[k].psi_in, x, c, cols, qh);
                h_fetch<T, SEAM>(g, a.hs[k], c, x, pbytes, in[k]);
                if (k == 1) fetch_e();
            } else if (valid) {
                psi_fetch_cols(g, a.hs[k].psi_in, x, c, cols, qh);
            }
            if (valid) {
                const HCell<T> h = h_update<T, SEAM, true>(g, a.hs[k], c, x, pbytes, q[k][0], q[k][1], q[k][2], in[k],
                                                           iu, iu, iu, own && !dry, (const CpmlCols<T, AX>*)&cols,
                                                           nullptr, &qh, tab);
                h1[k][0] = h.hx1; h1[k][1] = h.hy1; h1[k][2] = h.hz1;
                if (k == 0) {
                    const T aHx = h.hx1 + h.hx0, aHy = h.hy1 + h.hy0, aHz = h.hz1 + h.hz0;
                    oSxy = h.aEyx * aHz;
                    oSxz = -h.aEzx * aHy;
                    oSyz = h.aEzy * aHx;
                    oSyx = -h.aExy * aHz;
                    oSzx = h.aExz * aHy;
                    oSzy = -h.aEyz * aHx;
                }
            }
        }
        const int b = it & 1;
        lds[b][0][t] = oSxy;
        lds[b][1][t] = oSzy;
        lds[b][2][t] = oSxz;
        lds[b][3][t] = oSyz;
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int r = 0; r < 3; ++r) lds[b][4 + 3 * k + r][t] = h1[k][r];
        lds_barrier();
        // ---- S, E^{n+1} and F^{n+1} on the owned cells
        if (own && !dry) {
            const T nxy = lds[b][0][ty], nzy = lds[b][1][ty];
            const T nxz = lds[b][2][tz], nyz = lds[b][3][tz];
            stc<FDTD_NT_H>(a.S, 0, oi, (oSxy + oSxz + nxy + nxz) * sf);
            stc<FDTD_NT_H>(a.S, 1, oi, (oSyz + oSyx + nyz + oSyx_m) * sf);
            stc<FDTD_NT_H>(a.S, 2, oi, (oSzx + oSzy + oSzx_m + nzy) * sf);
            const bool src = src_comp >= 0 && x == a.src_x;
            const T amp = src ? *a.src_amp : (T)0;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int o = 4 + 3 * k, po = 4 + 3 * (1 - k);
                EIn<T> v;
                v.hx = h1[k][0]; v.hy = h1[k][1]; v.hz = h1[k][2];
                v.hx_ym = lds[b][o + 0][ty];
                v.hz_ym = lds[b][o + 2][ty];
                v.hx_zm = lds[b][o + 0][tz];
                v.hy_zm = lds[b][o + 1][tz];
                T hy_x = hym[k], hz_x = hzm[k];
                if constexpr (SEAM) {
                    if ((g.seam_mask & 2) && c.y == 0) {
                        v.hx_ym = seam_mix(v.hx_ym, lds[b][po + 0][ty], a.esa[k][1], a.esb[k][1]);
                        v.hz_ym = seam_mix(v.hz_ym, lds[b][po + 2][ty], a.esa[k][1], a.esb[k][1]);
                    }
                    if ((g.seam_mask & 4) && c.z == 0) {
                        v.hx_zm = seam_mix(v.hx_zm, lds[b][po + 0][tz], a.esa[k][2], a.esb[k][2]);
                        v.hy_zm = seam_mix(v.hy_zm, lds[b][po + 1][tz], a.esa[k][2], a.esb[k][2]);
                    }
                    if ((g.seam_mask & 1) && x == 0) {
                        hy_x = seam_mix(hym[k], hym[1 - k], a.esa[k][0], a.esb[k][0]);
                        hz_x = seam_mix(hzm[k], hzm[1 - k], a.esa[k][0], a.esb[k][0]);
                    }
                }
                v.ex = q[k][0]; v.ey = q[k][1]; v.ez = q[k][2];
                e_update<T, true>(g, a.es[k], c, x, oi, v, hy_x, hz_x, e0, e1, e2, iu, iu, iu,
                                  k == 0 && src ? src_comp : -1, amp, (const CpmlCols<T, AX>*)&cols, nullptr, &qe[k], tab);
            }
        }
        oSyx_m = oSyx;
        oSzx_m = oSzx;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            hym[k] = h1[k][1];
            hzm[k] = h1[k][2];
            q[k][0] = in[k].exn;
            q[k][1] = in[k].eyn;
            q[k][2] = in[k].ezn;
        }
        x = xn;
    }
}

// ---------------------------------------------------------------------------
// Two full steps in one march (he2_march): H^{n+1/2}, E^{n+1}, H^{n+3/2}
// with S, E^{n+2}.  E^n, H^{n-1/2}, eps and the psi blocks are read once and
// E^{n+2}, H^{n+3/2}, S and the psi blocks written once for the two steps;
// the fields of step n+1 never leave the chip.
//
// The march runs two stages one plane apart.  At plane x the first stage
// evaluates H^{n+1/2}(x) and E^{n+1}(x) as he_march does (H at y-1, z-1
// through LDS, at x-1 from the previous iteration); the second stage then
// evaluates H^{n+3/2}(x-1), which needs E^{n+1} at x-1 (previous
// iteration), at x (just computed) and at y+1, z+1 of plane x-1 (LDS,
// written one iteration before, kept by plane parity), and E^{n+2}(x-1) and
// S(x-1) from H^{n+3/2} through LDS.  The values of the previous iteration
// are read back from LDS (LC) or carried in registers.  Both E updates of a
// cell use one quotient dt / eps, divided once.  Each
// stage widens the work in y by one row, so a tile computes rows y0-2 .. y1
// for the owned rows y0 .. y1-1 (whole rows, z periodic inside a row), and
// an x chunk x0 .. x1-1 marches planes x0-2 .. x1.  The psi pairs of step
// n+1 stay in registers from the first stage to the second.  Everything the
// pass reads is read from buffers it does not write (E -> E_out, H -> H_out,
// psi_h -> psi_h_out, psi_e -> psi_e_out), because the halo rows and planes
// of one workgroup are owned cells of another.  Per cell the arithmetic is
// that of h_update + S and e_update, in the same order, so the results are
// bit-identical to two he_march passes (or to update_H, update_E, ...).
// S is stored for the second H half-step: the S of the first is overwritten
// by the same pass, at the same address and by the same thread, before
// anything can read it, so the pass leaves every buffer as two he_march
// passes do.  S2 stores it anyway (one more 12-byte store per cell; 6 %
// slower on a 395 x 345 x 73 grid).
// ---------------------------------------------------------------------------
template <class T>
struct HE2Args {
    Geom<T> g;
    Packed3<const T> E, H, eps;     // E^n, H^{n-1/2}
    Packed3<T> E_out, H_out, S;     // E^{n+2}, H^{n+3/2}, S
    const T* psi_h_in;
    T* psi_h_out;
    const T* psi_e_in;
    T* psi_e_out;
    const T* src_amp;               // two values: added to E^{n+1} and E^{n+2} at the source; null: none
    int src_x, src_p, src_comp;
    int coef_elems;
    // learned shell (SHELL): the correction added to E^n before the first
    // E update on the cells within ycT of a face, per component on the grid
    Packed3<const float> ycorr;
    int ycT;
};

// LC: H^{n+1/2}, E^{n+1} and eps of plane x-1 are read back from LDS (by
// plane parity) instead of being carried in registers (fewer spills).
// EARLY: the loads of plane x+1 are issued after the second barrier, so
// they are in flight while the second stage finishes plane x-1.
// POY: fused Poynting (S of the second H half-step); without it (and then
// without S2) the pass skips the oS products and S.
// SHELL: add the pending learned-shell correction to E^n before the first
// E update, on every cell that update evaluates (the halo rows' E^{n+1}
// feeds the second H half-step); the first H update and S use E^n as it is.
// NB: E^n at y+1, z+1 read from LDS (the tile's last row, and with z
// segments its last column, still load them from global memory).
template <class T, bool NU, bool S2, bool E2 = true, bool LC = (FDTD_HE2_LDSCARRY != 0),
          bool EARLY = (FDTD_HE2_EARLY != 0), bool POY = true, bool SHELL = false, bool NB = false>
__global__ __launch_bounds__(FDTD_HE2_HB, FDTD_HE2_HB >= 1024 ? FDTD_HE_WAVES : 3) void he2_march(HE2Args<T> a) {
    static_assert(POY || !S2, "the first half-step's S needs Poynting");
    constexpr int HB = FDTD_HE2_HB;
    __shared__ T la[LC ? 2 : 1][S2 ? 7 : 3][HB];   // H^{n+1/2} (x, y, z), with S2 also its oS products
    __shared__ T lp[LC ? 2 : 1][3][LC ? HB : 1];   // dt / eps
    __shared__ T le[2][3][HB];         // E^{n+1} by plane parity
    __shared__ T lb[7][HB];            // oS products of H^{n+3/2}, then H^{n+3/2} (x, y, z)
    __shared__ T ctab[kCoefTab];
    // NB: E^n of plane x+1 (loaded anyway for the first H update of plane x)
    // is put in LDS by plane parity, so the first H update of plane x+1 reads
    // its y+1 and z+1 neighbours there; only the tile's last row (and, with z
    // segments, its last column) loads them from global memory
    __shared__ T lx[NB ? 2 : 1][3][NB ? HB : 1];
    const Geom<T>& g = a.g;
    int pb, ck;
    block_coords(g.pblocks, g.nchunk, &pb, &ck);
    const int t = threadIdx.x;
    // rows y0-2 .. y1 for the owned rows y0 .. y1-1 (with z segments also
    // columns z0-2 .. z1 for the owned cells z0 .. z1-1)
    const TileCell tc = tile_cell<2, 1, 2, 1>(g, pb, t);
    const bool valid = tc.in;                                                   // H^{n+1/2}
    const bool rowe = valid && tc.r >= 1 && (!tc.seg || tc.cz >= 1);           // E^{n+1}: y0-1 .. y1
    const bool rowh = rowe && tc.r < tc.nr - 1 && (!tc.seg || tc.cz < tc.ncz - 1);   // H^{n+3/2}: .. y1-1
    const bool own = rowh && tc.r >= 2 && (!tc.seg || tc.cz >= 2);             // y0 .. y1-1
    const int p = tc.p;
    const Cell c = make_cell<T>(g, valid ? p : 0);
    const uint32_t pbytes = (uint32_t)g.plane * (uint32_t)sizeof(T);
    const int x0 = ck * g.xchunk;
    const int x1 = min(x0 + g.xchunk, g.sx);
    const int n_it = (x1 - x0) + 3;
    const T sf = g.dt * (T)0.125;
    const T iy = NU ? g.inv_e[1][c.y] : g.inv_u;     // H steps
    const T iz = NU ? g.inv_e[2][c.z] : g.inv_u;
    const T jy = NU ? g.inv_h[1][c.y] : g.inv_u;     // E steps
    const T jz = NU ? g.inv_h[2][c.z] : g.inv_u;
    const int tw = g.tw;
    const int tym = t - tw, typ = t + tw;
    const int tzm = (tc.cz > 0) ? t - 1 : t + tw - 1;
    const int tzp = (tc.cz < tw - 1) ? t + 1 : t - (tw - 1);
    const int src_comp = (valid && a.src_amp != nullptr && p == a.src_p) ? a.src_comp : -1;
    for (int i = t; i < a.coef_elems; i += HB) ctab[i] = g.coef[i];
    lds_barrier();
    const T* tab = ctab;
    CpmlCols<T> cols;
    cols.y = psi_col<1>(g, c, slab_hit(g, 1, c.y, g.sy));
    cols.z = psi_col<2>(g, c, slab_hit(g, 2, c.z, g.sz));
    HSysP<T> hA{a.E, a.H, {nullptr, 0u}, a.psi_h_in, nullptr};
    HSysP<T> hB{a.E, a.H, a.H_out, a.psi_h_in, a.psi_h_out};
    ESysP<T> eA{{nullptr, 0u}, (T*)a.psi_e_in};
    ESysP<T> eB{a.E_out, a.psi_e_out};

    int xa = (x0 - 2) % g.sx;
    if (xa < 0) xa += g.sx;
    T qx = (T)0, qy = (T)0, qz = (T)0;                  // E^n at plane xa
    if (valid) {
        const uint32_t ob = (uint32_t)xa * pbytes;
        qx = ldc<false>(a.E, 0, ob + c.op);
        qy = ldc<false>(a.E, 1, ob + c.op);
        qz = ldc<false>(a.E, 2, ob + c.op);
    }
    // NB: the neighbours a cell takes from global memory (outside the tile)
    const bool gy = tc.r == tc.nr - 1;
    const bool gz = tc.seg && tc.cz == tc.ncz - 1;
    if constexpr (NB) {
        lx[0][0][t] = qx;
        lx[0][1][t] = qy;
        lx[0][2][t] = qz;
        lds_barrier();
    }
    // carried from plane xa-1: H^{n+1/2}, E^{n+1}, the psi pairs after the
    // first half-steps, eps, (S2) oS of the first H half-step; from xa-2:
    // H^{n+3/2} and its oS products
    T h1x = (T)0, h1y = (T)0, h1z = (T)0;
    T f1x = (T)0, f1y = (T)0, f1z = (T)0;
    PsiPair<T> kh{}, ke{};
    T ep0 = (T)1, ep1 = (T)1, ep2 = (T)1;
    T oAyx_m = (T)0, oAzx_m = (T)0;
    T h2y_m = (T)0, h2z_m = (T)0, oByx_m = (T)0, oBzx_m = (T)0;
    // everything the first stage loads for plane xf (the E-step part when ef)
    struct Loads {
        HIn<T> h;
        PsiPair<T> qh, qe;    // (left unset where the cell is in no slab: only read where it is)
        T e0, e1, e2;
        bool ins;             // the cell is in the shell (SHELL)
        float y0, y1, y2;     // its pending correction
    };
    const bool in_yz = SHELL && shell_yz(c.y, c.z, a.ycT, g.sy, g.sz);
    const auto fetch = [&](int xf, bool ef, Loads& L) {
        CpmlCols<T> kf = cols;
        kf.hx = slab_hit(g, 0, xf, g.sx);
        L.e0 = L.e1 = L.e2 = (T)1;
        if (valid) psi_fetch_cols(g, a.psi_h_in, xf, c, kf, L.qh);
        if (rowe && ef) {
            const uint32_t of = (uint32_t)xf * pbytes + c.op;
            psi_fetch_cols(g, a.psi_e_in, xf, c, kf, L.qe);
            L.e0 = ldc<FDTD_NT_E>(a.eps, 0, of);
            L.e1 = ldc<FDTD_NT_E>(a.eps, 1, of);
            L.e2 = ldc<FDTD_NT_E>(a.eps, 2, of);
        }
        L.ins = false;
        if (SHELL && rowe && ef) {
            L.ins = xf < a.ycT || xf >= g.sx - a.ycT || in_yz;
            if (L.ins) {
                const uint32_t of = (uint32_t)xf * pbytes + c.op;
                L.y0 = ldc<false>(a.ycorr, 0, of);
                L.y1 = ldc<false>(a.ycorr, 1, of);
                L.y2 = ldc<false>(a.ycorr, 2, of);
            }
        }
        if constexpr (NB) {
            // h_fetch without the y+1 / z+1 neighbours inside the tile
            const uint32_t ob = (uint32_t)xf * pbytes;
            const uint32_t on = (xf == g.sx - 1) ? 0u : ob + pbytes;
            L.h.exn = ldc<FDTD_NT_H>(hA.E, 0, on + c.op);
            L.h.eyn = ldc<FDTD_NT_H>(hA.E, 1, on + c.op);
            L.h.ezn = ldc<FDTD_NT_H>(hA.E, 2, on + c.op);
            if (gy) {
                L.h.ex_y = ldc<false>(hA.E, 0, ob + c.op_y);
                L.h.ez_y = ldc<false>(hA.E, 2, ob + c.op_y);
            }
            if (gz) {
                L.h.ex_z = ldc<false>(hA.E, 0, ob + c.op_z);
                L.h.ey_z = ldc<false>(hA.E, 1, ob + c.op_z);
            }
            L.h.hx = ldc<FDTD_NT_H>(hA.Hin, 0, ob + c.op);
            L.h.hy = ldc<FDTD_NT_H>(hA.Hin, 1, ob + c.op);
            L.h.hz = ldc<FDTD_NT_H>(hA.Hin, 2, ob + c.op);
        } else {
            h_fetch<T, false>(g, hA, c, xf, pbytes, L.h);
        }
    };
    Loads cur;
    if (EARLY) fetch(xa, false, cur);
    for (int it = 0; it < n_it; ++it) {
        const int xb = (xa == 0) ? g.sx - 1 : xa - 1;
        const int xn = (xa == g.sx - 1) ? 0 : xa + 1;
        const bool ea = it >= 1;                 // E^{n+1} at xa
        const bool hb = it >= 2;                 // H^{n+3/2} at xb
        const bool wb = it >= 3;                 // xb is in the chunk: stores
        // ---- loads of plane xa (issued one iteration earlier with EARLY)
        CpmlCols<T> ka = cols, kb = cols;
        ka.hx = slab_hit(g, 0, xa, g.sx);
        kb.hx = slab_hit(g, 0, xb, g.sx);
        if (!EARLY) fetch(xa, ea, cur);
        if constexpr (NB) {
            // E^n of plane xa at y+1 and z+1, written one iteration earlier
            // (or before the march) into the slot of this iteration's parity
            const int s0 = it & 1;
            if (!gy) {
                cur.h.ex_y = lx[s0][0][min(typ, HB - 1)];
                cur.h.ez_y = lx[s0][2][min(typ, HB - 1)];
            }
            if (!gz) {
                cur.h.ex_z = lx[s0][0][tzp];
                cur.h.ey_z = lx[s0][1][tzp];
            }
        }
        const HIn<T>& in = cur.h;
        const PsiPair<T>& qh = cur.qh;
        const PsiPair<T>& qe = cur.qe;
        // dt / eps of plane xa, divided once for both E half-steps (the
        // same IEEE quotient the E update forms)
        const T e0 = g.dt / cur.e0, e1 = g.dt / cur.e1, e2 = g.dt / cur.e2;
        const int b = it & 1;
        const int ba = LC ? b : 0;
        // ---- first stage: H^{n+1/2}(xa) on every row
        const T ixa = NU ? g.inv_e[0][xa] : g.inv_u;
        PsiPair<T> nkh, nke;
        T g1x = (T)0, g1y = (T)0, g1z = (T)0;
        T oAxy = (T)0, oAxz = (T)0, oAyz = (T)0, oAyx = (T)0, oAzx = (T)0, oAzy = (T)0;
        if (valid) {
            const HCell<T> h = h_update<T, false, true>(g, hA, c, xa, pbytes, qx, qy, qz, in, ixa, iy, iz, false,
                                                        (const CpmlCols<T>*)&ka, nullptr, &qh, tab, &nkh);
            g1x = h.hx1; g1y = h.hy1; g1z = h.hz1;
            if (S2) {
                const T aHx = h.hx1 + h.hx0, aHy = h.hy1 + h.hy0, aHz = h.hz1 + h.hz0;
                oAxy = h.aEyx * aHz;
                oAxz = -h.aEzx * aHy;
                oAyz = h.aEzy * aHx;
                oAyx = -h.aExy * aHz;
                oAzx = h.aExz * aHy;
                oAzy = -h.aEyz * aHx;
            }
        }
        la[ba][0][t] = g1x;
        la[ba][1][t] = g1y;
        la[ba][2][t] = g1z;
        if constexpr (NB) {
            // E^n of plane xa + 1 for the next iteration (its slot was last
            // read before the previous iteration's first barrier)
            lx[(it + 1) & 1][0][t] = in.exn;
            lx[(it + 1) & 1][1][t] = in.eyn;
            lx[(it + 1) & 1][2][t] = in.ezn;
        }
        if (S2) {
            la[ba][3][t] = oAxy;
            la[ba][4][t] = oAzy;
            la[ba][5][t] = oAxz;
            la[ba][6][t] = oAyz;
        }
        lds_barrier();
        // ---- first stage: S(xa) (S2) and E^{n+1}(xa) on rows y0-1 .. y1
        T n1x = (T)0, n1y = (T)0, n1z = (T)0;
        if (rowe && ea) {
            if (S2 && own && hb && it < n_it - 1) {
                const uint32_t oi = (uint32_t)xa * pbytes + c.op;
                stc<FDTD_NT_H>(a.S, 0, oi, (oAxy + oAxz + la[ba][3][tym] + la[ba][5][tzm]) * sf);
                stc<FDTD_NT_H>(a.S, 1, oi, (oAyz + oAyx + la[ba][6][tzm] + oAyx_m) * sf);
                stc<FDTD_NT_H>(a.S, 2, oi, (oAzx + oAzy + oAzx_m + la[ba][4][tym]) * sf);
            }
            EIn<T> v;
            v.hx = g1x; v.hy = g1y; v.hz = g1z;
            v.hx_ym = la[ba][0][tym];
            v.hz_ym = la[ba][2][tym];
            v.hx_zm = la[ba][0][tzm];
            v.hy_zm = la[ba][1][tzm];
            v.ex = qx; v.ey = qy; v.ez = qz;
            if (SHELL && cur.ins) {
                v.ex = v.ex + (T)cur.y0;
                v.ey = v.ey + (T)cur.y1;
                v.ez = v.ez + (T)cur.y2;
            }
            const T jxa = NU ? g.inv_h[0][xa] : g.inv_u;
            const bool src = src_comp >= 0 && xa == a.src_x;
            if constexpr (LC) {
                h1y = la[b ^ 1][1][t];
                h1z = la[b ^ 1][2][t];
            }
            // without E2 the pass ends after H^{n+3/2}: E^{n+1} and its psi
            // pairs are the pass's E output (owned cells of the chunk's planes)
            const bool wr1 = !E2 && own && it >= 2 && it < n_it - 1;
            e_value<T, true, ESysP<T>, const CpmlCols<T>*, true>(
                g, E2 ? eA : eB, c, xa, v, h1y, h1z, e0, e1, e2, jxa, jy, jz, src ? src_comp : -1,
                src ? a.src_amp[0] : (T)0, (const CpmlCols<T>*)&ka, nullptr, &qe, tab, wr1, &nke, n1x, n1y, n1z);
            if (wr1) {
                const uint32_t o1 = (uint32_t)xa * pbytes + c.op;
                stc<FDTD_NT_E>(eB.E, 0, o1, n1x);
                stc<FDTD_NT_E>(eB.E, 1, o1, n1y);
                stc<FDTD_NT_E>(eB.E, 2, o1, n1z);
            }
        }
        le[b][0][t] = n1x;
        le[b][1][t] = n1y;
        le[b][2][t] = n1z;
        if constexpr (LC) {
            lp[b][0][t] = e0;
            lp[b][1][t] = e1;
            lp[b][2][t] = e2;
        }
        // ---- second stage: H^{n+3/2}(xb) on rows y0-1 .. y1-1
        T g2x = (T)0, g2y = (T)0, g2z = (T)0;
        T oBxy = (T)0, oBxz = (T)0, oByz = (T)0, oByx = (T)0, oBzx = (T)0, oBzy = (T)0;
        if (rowh && hb) {
            HIn<T> v;
            v.exn = n1x; v.eyn = n1y; v.ezn = n1z;
            v.ex_y = le[b ^ 1][0][typ];
            v.ez_y = le[b ^ 1][2][typ];
            v.ex_z = le[b ^ 1][0][tzp];
            v.ey_z = le[b ^ 1][1][tzp];
            if constexpr (LC) {
                h1x = la[b ^ 1][0][t];
                h1y = la[b ^ 1][1][t];
                h1z = la[b ^ 1][2][t];
                f1x = le[b ^ 1][0][t];
                f1y = le[b ^ 1][1][t];
                f1z = le[b ^ 1][2][t];
            }
            v.hx = h1x; v.hy = h1y; v.hz = h1z;
            const T ixb = NU ? g.inv_e[0][xb] : g.inv_u;
            const HCell<T> h = h_update<T, false, true>(g, hB, c, xb, pbytes, f1x, f1y, f1z, v, ixb, iy, iz,
                                                        own && wb, (const CpmlCols<T>*)&kb, nullptr, &kh, tab);
            if constexpr (POY) {
                const T aHx = h.hx1 + h.hx0, aHy = h.hy1 + h.hy0, aHz = h.hz1 + h.hz0;
                oBxy = h.aEyx * aHz;
                oBxz = -h.aEzx * aHy;
                oByz = h.aEzy * aHx;
                oByx = -h.aExy * aHz;
                oBzx = h.aExz * aHy;
                oBzy = -h.aEyz * aHx;
            }
            g2x = h.hx1; g2y = h.hy1; g2z = h.hz1;
        }
        if constexpr (POY) {
            lb[0][t] = oBxy;
            lb[1][t] = oBzy;
            lb[2][t] = oBxz;
            lb[3][t] = oByz;
        }
        lb[4][t] = g2x;
        lb[5][t] = g2y;
        lb[6][t] = g2z;
        lds_barrier();
        Loads nxt;
        if (EARLY && it + 1 < n_it) fetch(xn, true, nxt);
        // ---- second stage: S(xb) and E^{n+2}(xb) on the owned rows
        if (POY && own && wb) {
            const uint32_t oi = (uint32_t)xb * pbytes + c.op;
            stc<FDTD_NT_H>(a.S, 0, oi, (oBxy + oBxz + lb[0][tym] + lb[2][tzm]) * sf);
            stc<FDTD_NT_H>(a.S, 1, oi, (oByz + oByx + lb[3][tzm] + oByx_m) * sf);
            stc<FDTD_NT_H>(a.S, 2, oi, (oBzx + oBzy + oBzx_m + lb[1][tym]) * sf);
        }
        if (E2 && own && wb) {
            const uint32_t oi = (uint32_t)xb * pbytes + c.op;
            EIn<T> v;
            v.hx = g2x; v.hy = g2y; v.hz = g2z;
            v.hx_ym = lb[4][tym];
            v.hz_ym = lb[6][tym];
            v.hx_zm = lb[4][tzm];
            v.hy_zm = lb[5][tzm];
            if constexpr (LC) {
                v.ex = le[b ^ 1][0][t];
                v.ey = le[b ^ 1][1][t];
                v.ez = le[b ^ 1][2][t];
                ep0 = lp[b ^ 1][0][t];
                ep1 = lp[b ^ 1][1][t];
                ep2 = lp[b ^ 1][2][t];
            } else {
                v.ex = f1x; v.ey = f1y; v.ez = f1z;
            }
            const T jxb = NU ? g.inv_h[0][xb] : g.inv_u;
            const bool src = src_comp >= 0 && xb == a.src_x;
            T ox, oy, oz;
            e_value<T, true, ESysP<T>, const CpmlCols<T>*, true>(
                g, eB, c, xb, v, h2y_m, h2z_m, ep0, ep1, ep2, jxb, jy, jz, src ? src_comp : -1,
                src ? a.src_amp[1] : (T)0, (const CpmlCols<T>*)&kb, nullptr, &ke, tab, true, (PsiPair<T>*)nullptr, ox,
                oy, oz);
            stc<FDTD_NT_E>(eB.E, 0, oi, ox);
            stc<FDTD_NT_E>(eB.E, 1, oi, oy);
            stc<FDTD_NT_E>(eB.E, 2, oi, oz);
        }
        // ---- carry to the next plane
        h2y_m = g2y; h2z_m = g2z;
        oByx_m = oByx; oBzx_m = oBzx;
        if constexpr (!LC) {
            h1x = g1x; h1y = g1y; h1z = g1z;
            f1x = n1x; f1y = n1y; f1z = n1z;
            ep0 = e0; ep1 = e1; ep2 = e2;
        }
        kh = nkh;
        ke = nke;
        if (S2) { oAyx_m = oAyx; oAzx_m = oAzx; }
        qx = in.exn; qy = in.eyn; qz = in.ezn;
        if (EARLY) cur = nxt;
        xa = xn;
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
    const Cell c = make_cell<T>(g, p);
    const uint32_t pbytes = (uint32_t)g.plane * (uint32_t)sizeof(T);
    const uint32_t ob = (uint32_t)x * pbytes;
    const uint32_t i = ob + c.op;
    const uint32_t imy = ob + c.om_y;
    const uint32_t imz = ob + c.om_z;
    const uint32_t imx = (uint32_t)((x == 0) ? g.sx - 1 : x - 1) * pbytes + c.op;
    const T sf = g.dt * (T)0.125;
    st(a.S[0], i, (ld(a.oS[0], i) + ld(a.oS[3], i) + ld(a.oS[0], imy) + ld(a.oS[3], imz)) * sf);
    st(a.S[1], i, (ld(a.oS[1], i) + ld(a.oS[4], i) + ld(a.oS[1], imz) + ld(a.oS[4], imx)) * sf);
    st(a.S[2], i, (ld(a.oS[2], i) + ld(a.oS[5], i) + ld(a.oS[2], imx) + ld(a.oS[5], imy)) * sf);
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

#ifdef FDTD_RECORD_LAUNCHES
#include "launch_record.inl"
#endif

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// A marching launch: block threads per workgroup, pblocks tiles of the
// plane times nchunk x chunks of xchunk planes.  tile: the plane cells a
// workgroup owns.  he_march / he2_march tiles (tyo, tzo, tzb, tw, as in
// Geom): whole rows (tzb = 1, tzo = tw = sz), or rows cut into z segments
// of tzo owned cells, each workgroup row then tw = tzo + halo columns wide.
struct Launch {
    int block, tile, pblocks, nchunk, xchunk, mode;
    int tyo = 0, tzo = 0, tzb = 0, tw = 0;
};

struct Plan {
    fdtd_desc d;
    int sx, sy, sz, plane;
    int64_t N;
    Launch le, lh;
    Launch lhe;          // fused H+E pass (mode 2 when available, else mode -1)
    Launch lhe2;         // two-step pass (mode 2 when available, else mode -1)
    bool he2_pref;       // step_n runs the two-step pass (its cost model beats the fused pass's)
    int64_t psi_elems, coef_elems;
    struct F { int lo, hi, psi_off, face_elems, coef_off; } face[FDTD_MAX_FACES];
};

int64_t plane_of(const Plan& P, int axis) {
    if (axis == 0) return (int64_t)P.sy * P.sz;
    if (axis == 1) return (int64_t)P.sx * P.sz;
    return (int64_t)P.sx * P.sy;
}

Launch make_launch(int plane, int sx, int block, int tile, int mode, int target) {
    Launch L;
    L.block = block;
    L.tile = tile;
    L.mode = mode;
    L.pblocks = (plane + tile - 1) / tile;
    int nchunk = (target + L.pblocks - 1) / L.pblocks;
    if (nchunk > sx) nchunk = sx;
    if (nchunk < 1) nchunk = 1;
    L.xchunk = (sx + nchunk - 1) / nchunk;
    L.nchunk = (sx + L.xchunk - 1) / L.xchunk;
    return L;
}

// The time model of a marching launch with one workgroup per CU at a time:
// each round runs up to max_blocks workgroups side by side, and a chunk
// takes a time proportional to its planes, dry ones included (extra).  For
// pblocks tiles per plane, the number of x chunks per tile column that
// minimises rounds x (planes per chunk + extra): returns that cost, the
// chunk count in *best_n.
int64_t round_chunks(int pblocks, int sx, int max_blocks, int extra, int* best_n) {
    int best = 1;
    int64_t best_cost = -1;
    for (int n = 1; n <= sx; ++n) {
        const int xchunk = (sx + n - 1) / n;
        const int nchunk = (sx + xchunk - 1) / xchunk;
        if (nchunk != n) continue;                       // same chunking as a smaller n
        const int64_t blocks = (int64_t)pblocks * nchunk;
        const int64_t rounds = (blocks + max_blocks - 1) / max_blocks;
        const int64_t cost = rounds * (xchunk + extra);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = n;
        }
    }
    *best_n = best;
    return best_cost;
}

// Marching launch of tiles of `tile` contiguous plane cells (whole rows
// plus a halo row for the fused-S H kernel), chunked by that model.
Launch make_round_launch(int plane, int sx, int block, int tile, int max_blocks, int extra = 1) {
    const int pblocks = (plane + tile - 1) / tile;
    int best = 1;
    (void)round_chunks(pblocks, sx, max_blocks, extra, &best);
    return make_launch(plane, sx, block, tile, 2, best * pblocks);
}

// Launch of he_march / he2_march: tiles of tyo rows by tzo z cells, the
// workgroup holding yh halo rows besides the owned rows and, with z
// segments, zh halo columns besides the owned cells.  Candidates: whole
// rows where 1 + yh of them fit the workgroup (z periodic inside a row),
// and unless whole_only, z segments of FDTD_SEG_MIN (64) up to 256 owned
// cells, powers of two, on longer rows.  Each is chunked along x by
// round_chunks; the one with the fewest rounds x planes wins, whole rows
// on a tie.  Cutting rows
// keeps long rows from leaving a tile mostly halo (rows of 512 cells: one
// owned row of two in a 1024-thread workgroup).  mode -1: no shape fits.
Launch make_tile_launch(int sx, int sy, int sz, int block, int yh, int zh, int max_blocks, int extra,
                        bool whole_only) {
    Launch best{};
    best.mode = -1;
    int64_t best_cost = -1;
    const auto consider = [&](int tyo, int tzo, int tzb, int tw) {
        if (tyo < 1) return;
        const int pblocks = ((sy + tyo - 1) / tyo) * tzb;
        int n = 1;
        const int64_t cost = round_chunks(pblocks, sx, max_blocks, extra, &n);
        if (best_cost >= 0 && cost >= best_cost) return;
        best_cost = cost;
        best.block = block;
        best.mode = 2;
        best.tyo = tyo;
        best.tzo = tzo;
        best.tzb = tzb;
        best.tw = tw;
        best.tile = tyo * tzo;
        best.pblocks = pblocks;
        best.xchunk = (sx + n - 1) / n;
        best.nchunk = (sx + best.xchunk - 1) / best.xchunk;
    };
    if ((int64_t)(1 + yh) * sz <= block) consider(block / sz - yh, sz, 1, sz);
    if (!whole_only) {
        for (int tzo = FDTD_SEG_MIN; tzo <= 256; tzo *= 2)
            if (tzo < sz && tzo + zh <= block) consider(block / (tzo + zh) - yh, tzo, (sz + tzo - 1) / tzo, tzo + zh);
        // rows cut into k equal segments (rows of 136: two of 68, 11 owned
        // rows per tile instead of 4 whole rows or 12 rows of 64 + 64 + 8)
        for (int k = 2; k <= 16; ++k) {
            const int tzo = (sz + k - 1) / k;
            if (tzo < FDTD_SEG_MIN) break;
            if (tzo < sz && tzo <= 256 && tzo + zh <= block)
                consider(block / (tzo + zh) - yh, tzo, (sz + tzo - 1) / tzo, tzo + zh);
        }
    }
    return best;
}

template <class T>
void fill_geom(const Plan& P, const Launch& L, const fdtd_fields* f, Geom<T>& g) {
    g.sx = P.sx; g.sy = P.sy; g.sz = P.sz; g.plane = P.plane;
    g.xchunk = L.xchunk;
    g.nchunk = L.nchunk;
    g.pblocks = L.pblocks;
    g.tile = L.tile;
    g.tyo = L.tyo;
    g.tzo = L.tzo;
    g.tzb = L.tzb;
    g.tw = L.tw;
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
        fd.psi_off = P.face[k].psi_off;
        fd.face_elems = P.face[k].face_elems;
        fd.coef_off = P.face[k].coef_off;
    }
}

template <class T>
void comp3(const void* base, int64_t N, T* out[3]) {
    for (int k = 0; k < 3; ++k) out[k] = base ? (T*)base + k * N : nullptr;
}
template <class T>
void comp3c(const void* base, int64_t N, const T* out[3]) {
    for (int k = 0; k < 3; ++k) out[k] = base ? (const T*)base + k * N : nullptr;
}

int check_launch(const char* what) {
#ifdef FDTD_RECORD_LAUNCHES
    (void)what;
    return 0;     // nothing was launched
#endif
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(5, "%s: launch failed: %s", what, hipGetErrorString(e));
    return 0;
}

template <class T>
int launch_e(const Plan& P, const fdtd_fields* f, hipStream_t st, const T* src_amp, int64_t src_index,
             const float* ycorr = nullptr, int ycT = 0, int64_t ycells = 0) {
    EArgs<T> a;
    memset(&a, 0, sizeof(a));
    if (ycorr) {
        for (int r = 0; r < 3; ++r) a.ycorr[r] = ycorr + (size_t)r * ycells;
        a.ycT = ycT;
    }
    fill_geom<T>(P, P.le, f, a.g);
    const int64_t N = P.N;
    comp3c<T>(f->eps, N, a.eps);
    const bool pair = P.d.bloch_mask != 0;
    const double* re = P.d.bloch_re;
    const double* im = P.d.bloch_im;
    // system 0: E from H, partner G; seam H(-1) = (H + iG)(c - is) -> real: Hc + Gs
    comp3<T>(f->E, N, a.s[0].E);
    comp3c<T>(f->H, N, a.s[0].H);
    comp3c<T>(pair ? f->G : nullptr, N, a.s[0].P);
    a.s[0].psi = (T*)f->psi_e;
    for (int r = 0; r < 3; ++r) { a.s[0].sa[r] = (T)re[r]; a.s[0].sb[r] = (T)im[r]; }
    if (pair) {
        // system 1: F from G, partner H; imaginary part: Gc - Hs
        comp3<T>(f->F, N, a.s[1].E);
        comp3c<T>(f->G, N, a.s[1].H);
        comp3c<T>(f->H, N, a.s[1].P);
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
    const dim3 grid(P.le.pblocks * P.le.nchunk);
    const bool seam = pair && P.d.bloch_mode == FDTD_BLOCH_PHASE;
    if (ycorr && pair) return fail(3, "the shell correction is for real fields (no Bloch boundaries)");
    a.coef_elems = (int)P.coef_elems;
    const bool ct = FDTD_EH_CTAB && P.coef_elems <= kCoefTab;
#define FDTD_LAUNCH_E(PA, SE, SH)                                                                            \
    do {                                                                                                     \
        if (ct) hipLaunchKernelGGL((e_march<T, PA, SE, SH, true>), grid, dim3(kBlock), 0, st, a);            \
        else hipLaunchKernelGGL((e_march<T, PA, SE, SH, false>), grid, dim3(kBlock), 0, st, a);             \
    } while (0)
    if (ycorr) FDTD_LAUNCH_E(false, false, true);
    else if (!pair) FDTD_LAUNCH_E(false, false, false);
    else if (!seam) FDTD_LAUNCH_E(true, false, false);
    else FDTD_LAUNCH_E(true, true, false);
#undef FDTD_LAUNCH_E
    return check_launch("fdtd_update_e");
}

template <class T, bool PAIR, bool SEAM, bool CT>
void launch_h_mode(const Launch& L, hipStream_t st, const HArgs<T>& a) {
    const dim3 grid(L.pblocks * L.nchunk);
    switch (L.mode) {
        case 0: hipLaunchKernelGGL((h_march<T, PAIR, SEAM, 0, kBlock, CT>), grid, dim3(kBlock), 0, st, a); break;
        case 1: hipLaunchKernelGGL((h_march<T, PAIR, SEAM, 1, kBlock, CT>), grid, dim3(kBlock), 0, st, a); break;
        case 2:
            if (L.block == FDTD_HB_SMALL && FDTD_HB_SMALL != 1024)
                hipLaunchKernelGGL((h_march<T, PAIR, SEAM, 2, FDTD_HB_SMALL, CT>), grid, dim3(FDTD_HB_SMALL), 0, st,
                                   a);
            else hipLaunchKernelGGL((h_march<T, PAIR, SEAM, 2, 1024, CT>), grid, dim3(1024), 0, st, a);
            break;
        default: hipLaunchKernelGGL((h_march<T, PAIR, SEAM, 3, kBlock>), grid, dim3(kBlock), 0, st, a); break;
    }
}

template <class T>
void fill_h(const Plan& P, const Launch& L, const fdtd_fields* f, HArgs<T>& a) {
    memset(&a, 0, sizeof(a));
    fill_geom<T>(P, L, f, a.g);
    const int64_t N = P.N;
    const bool pair = P.d.bloch_mask != 0;
    const double* re = P.d.bloch_re;
    const double* im = P.d.bloch_im;
    // system 0: H from E, partner F; seam E(s) = (E + iF)(c + is) -> real: Ec - Fs
    comp3c<T>(f->E, N, a.s[0].E);
    comp3c<T>(f->H, N, a.s[0].Hin);
    comp3<T>(f->H_out ? f->H_out : f->H, N, a.s[0].Hout);
    comp3c<T>(pair ? f->F : nullptr, N, a.s[0].P);
    a.s[0].psi_in = (const T*)f->psi_h;
    a.s[0].psi_out = (T*)(f->psi_h_out ? f->psi_h_out : f->psi_h);
    for (int r = 0; r < 3; ++r) { a.s[0].sa[r] = (T)re[r]; a.s[0].sb[r] = (T)(-im[r]); }
    if (pair) {
        // system 1: G from F, partner E; imaginary part: Fc + Es
        comp3c<T>(f->F, N, a.s[1].E);
        comp3c<T>(f->G, N, a.s[1].Hin);
        comp3<T>(f->G_out ? f->G_out : f->G, N, a.s[1].Hout);
        comp3c<T>(f->E, N, a.s[1].P);
        a.s[1].psi_in = (const T*)f->psi_g;
        a.s[1].psi_out = (T*)(f->psi_g_out ? f->psi_g_out : f->psi_g);
        for (int r = 0; r < 3; ++r) { a.s[1].sa[r] = (T)re[r]; a.s[1].sb[r] = (T)im[r]; }
    }
    comp3<T>(f->S, N, a.S);
    for (int k = 0; k < 6; ++k) a.oS[k] = f->oS ? (T*)f->oS + k * N : nullptr;
}

template <class T>
int launch_h(const Plan& P, const fdtd_fields* f, hipStream_t st) {
    HArgs<T> a;
    fill_h<T>(P, P.lh, f, a);
    const bool pair = P.d.bloch_mask != 0;
    const bool seam = pair && P.d.bloch_mode == FDTD_BLOCH_PHASE;
    a.coef_elems = (int)P.coef_elems;
    // (the paired Bloch systems use it in mode 2 only, whose 1024-thread
    // tiles allow 128 VGPRs; in modes 0/1 the prefetched psi pairs spill)
    const bool ct = FDTD_EH_CTAB && P.coef_elems <= kCoefTab && (P.lh.mode == 2 || (P.lh.mode != 3 && !pair));
    if (!pair) ct ? launch_h_mode<T, false, false, true>(P.lh, st, a) : launch_h_mode<T, false, false, false>(P.lh, st, a);
    else if (!seam) ct ? launch_h_mode<T, true, false, true>(P.lh, st, a) : launch_h_mode<T, true, false, false>(P.lh, st, a);
    else ct ? launch_h_mode<T, true, true, true>(P.lh, st, a) : launch_h_mode<T, true, true, false>(P.lh, st, a);
    return check_launch("fdtd_update_h");
}

// (lhe is set up only where the pass applies: fused Poynting or Poynting
// off, 3N values within 32-bit byte offsets, and, with Bloch boundaries,
// fused Poynting, rows of at most FDTD_ROW_MAX cells, FP32, uniform dx and
// the coefficient block in LDS)
bool he_supported(const Plan& P) {
    return P.d.poynting != FDTD_POYNTING_TWO_PASS && P.lhe.mode == 2;
}

// The second buffers the fused passes read from and write to: E_out, H_out
// and (with CPML) psi_h_out.
bool he_buffers(const Plan& P, const fdtd_fields* f) {
    return f->E_out && f->E_out != f->E && f->H_out && f->H_out != f->H &&
           (P.d.n_face == 0 || (f->psi_h_out && f->psi_h_out != f->psi_h));
}

int launch_heb(const Plan& P, const fdtd_fields* f, hipStream_t st, const float* src_amp, int64_t src_index) {
    using T = float;
    HEBArgs<T> a;
    memset(&a, 0, sizeof(a));
    fill_geom<T>(P, P.lhe, f, a.g);
    const uint32_t nb = (uint32_t)(P.N * (int64_t)sizeof(T));
    const double* re = P.d.bloch_re;
    const double* im = P.d.bloch_im;
    // H from E (partner F): seam E(s) = (E + iF)(c + is) -> Ec - Fs; G from F: Fc + Es
    a.hs[0] = {{(const T*)f->E, nb}, {(const T*)f->H, nb}, {(const T*)f->F, nb}, {(T*)f->H_out, nb},
               (const T*)f->psi_h, (T*)f->psi_h_out, {}, {}};
    a.hs[1] = {{(const T*)f->F, nb}, {(const T*)f->G, nb}, {(const T*)f->E, nb}, {(T*)f->G_out, nb},
               (const T*)f->psi_g, (T*)f->psi_g_out, {}, {}};
    for (int r = 0; r < 3; ++r) {
        a.hs[0].sa[r] = (T)re[r]; a.hs[0].sb[r] = (T)(-im[r]);
        a.hs[1].sa[r] = (T)re[r]; a.hs[1].sb[r] = (T)im[r];
        // E from H (partner G): seam H(-1) = (H + iG)(c - is) -> Hc + Gs; F from G: Gc - Hs
        a.esa[0][r] = (T)re[r]; a.esb[0][r] = (T)im[r];
        a.esa[1][r] = (T)re[r]; a.esb[1][r] = (T)(-im[r]);
    }
    a.es[0] = {{(T*)f->E_out, nb}, (T*)f->psi_e};
    a.es[1] = {{(T*)f->F_out, nb}, (T*)f->psi_f};
    a.S = {(T*)f->S, nb};
    a.eps = {(const T*)f->eps, nb};
    a.coef_elems = (int)P.coef_elems;
    a.src_amp = src_amp;
    if (src_amp) {
        const int64_t N = P.N;
        const int64_t comp = src_index / N;
        const int64_t i = src_index - comp * N;
        a.src_comp = (int)comp;
        a.src_x = (int)(i / P.plane);
        a.src_p = (int)(i - (int64_t)a.src_x * P.plane);
    }
    const dim3 grid(P.lhe.pblocks * P.lhe.nchunk);
    // the axes that carry CPML faces: none, z only (the usual setup with
    // Bloch boundaries on x and y), or any
    int axes = 0;
    for (int k = 0; k < P.d.n_face; ++k) axes |= 1 << P.d.face[k].axis;
    const bool seam = P.d.bloch_mode == FDTD_BLOCH_PHASE;
#define FDTD_LAUNCH_HEB(AX_)                                                                                  \
    do {                                                                                                      \
        if (seam) hipLaunchKernelGGL((heb_march<T, true, AX_>), grid, dim3(1024), 0, st, a);                    \
        else hipLaunchKernelGGL((heb_march<T, false, AX_>), grid, dim3(1024), 0, st, a);                        \
    } while (0)
    if (axes == 0) FDTD_LAUNCH_HEB(0);
    else if (axes == 4) FDTD_LAUNCH_HEB(4);
    else FDTD_LAUNCH_HEB(7);
#undef FDTD_LAUNCH_HEB
    return check_launch("fdtd_update_he");
}

template <class T>
int launch_he(const Plan& P, const fdtd_fields* f, hipStream_t st, const T* src_amp, int64_t src_index,
              const float* ycorr = nullptr, int ycT = 0, int64_t ycells = 0) {
    if (P.d.bloch_mask) {
        if (sizeof(T) != 4 || ycorr) return fail(3, "the Bloch-pair pass is FP32 and takes no shell correction");
        return launch_heb(P, f, st, (const float*)src_amp, src_index);
    }
    HEArgs<T> a;
    memset(&a, 0, sizeof(a));
    if (ycorr) {
        a.ycorr = {ycorr, (uint32_t)(ycells * 4)};
        a.ycT = ycT;
    }
    fill_geom<T>(P, P.lhe, f, a.g);
    const uint32_t nb = (uint32_t)(P.N * (int64_t)sizeof(T));
    a.hs.E = {(const T*)f->E, nb};
    a.hs.Hin = {(const T*)f->H, nb};
    a.hs.Hout = {(T*)f->H_out, nb};
    a.hs.psi_in = (const T*)f->psi_h;
    a.hs.psi_out = (T*)f->psi_h_out;
    a.es.E = {(T*)f->E_out, nb};
    a.es.psi = (T*)f->psi_e;
    a.S = {(T*)f->S, nb};
    a.eps = {(const T*)f->eps, nb};
    const int64_t N = P.N;
    a.coef_elems = (int)P.coef_elems;
    a.src_amp = src_amp;
    if (src_amp) {
        const int64_t comp = src_index / N;
        const int64_t i = src_index - comp * N;
        a.src_comp = (int)comp;
        a.src_x = (int)(i / P.plane);
        a.src_p = (int)(i - (int64_t)a.src_x * P.plane);
    }
    const Launch& L = P.lhe;
    const dim3 grid(L.pblocks * L.nchunk);
    const bool nu = P.d.nonuniform != 0;
    const bool ct = FDTD_HE_CTAB && P.coef_elems <= kCoefTab;
    const bool poy = P.d.poynting == FDTD_POYNTING_FUSED;
#define FDTD_LAUNCH_HE3(HB_, SH_, POY_)                                                                      \
    do {                                                                                                     \
        if (nu && ct) hipLaunchKernelGGL((he_march<T, HB_, true, true, SH_, POY_>), grid, dim3(HB_), 0, st, a); \
        else if (nu) hipLaunchKernelGGL((he_march<T, HB_, true, false, SH_, POY_>), grid, dim3(HB_), 0, st, a); \
        else if (ct) hipLaunchKernelGGL((he_march<T, HB_, false, true, SH_, POY_>), grid, dim3(HB_), 0, st, a); \
        else hipLaunchKernelGGL((he_march<T, HB_, false, false, SH_, POY_>), grid, dim3(HB_), 0, st, a);        \
    } while (0)
#define FDTD_LAUNCH_HE2(HB_, SH_)                      \
    do {                                               \
        if (poy) FDTD_LAUNCH_HE3(HB_, SH_, true);      \
        else FDTD_LAUNCH_HE3(HB_, SH_, false);         \
    } while (0)
#define FDTD_LAUNCH_HE(HB_)                      \
    do {                                         \
        if (ycorr) FDTD_LAUNCH_HE2(HB_, true);   \
        else FDTD_LAUNCH_HE2(HB_, false);        \
    } while (0)
    if (L.block == FDTD_HE_HB_SMALL && FDTD_HE_HB_SMALL != 1024) FDTD_LAUNCH_HE(FDTD_HE_HB_SMALL);
    else FDTD_LAUNCH_HE(1024);
#undef FDTD_LAUNCH_HE
#undef FDTD_LAUNCH_HE2
#undef FDTD_LAUNCH_HE3
    return check_launch("fdtd_update_he");
}

int check_he_fields(const Plan& P, const fdtd_fields* f) {
    if (!he_supported(P))
        return fail(3, "fdtd_update_he needs fused Poynting or Poynting off and, with Bloch boundaries, fused "
                       "Poynting, rows of at most %d cells, FP32, uniform dx and at most %d CPML coefficients",
                    FDTD_ROW_MAX, kCoefTab);
    if (!f->E_out || f->E_out == f->E) return fail(2, "fdtd_update_he needs E_out != E");
    if (!f->H_out || f->H_out == f->H) return fail(2, "fdtd_update_he needs H_out != H");
    if (P.d.n_face > 0 && (!f->psi_h_out || f->psi_h_out == f->psi_h))
        return fail(2, "fdtd_update_he needs psi_h_out != psi_h");
    if (P.d.bloch_mask) {
        if (!f->F_out || f->F_out == f->F) return fail(2, "fdtd_update_he with Bloch boundaries needs F_out != F");
        if (!f->G_out || f->G_out == f->G) return fail(2, "fdtd_update_he with Bloch boundaries needs G_out != G");
        if (P.d.n_face > 0 && (!f->psi_g_out || f->psi_g_out == f->psi_g))
            return fail(2, "fdtd_update_he with Bloch boundaries needs psi_g_out != psi_g");
    }
    return 0;
}

bool he2_supported(const Plan& P) { return P.lhe2.mode == 2; }

// last_e false: the pass stops after the second H half-step (H, E with the
// source, H with S), which ends a step_n schedule with an odd pair count
int launch_he2(const Plan& P, const fdtd_fields* f, hipStream_t st, const float* src_amp, int64_t src_index,
               bool last_e = true, const float* ycorr = nullptr, int ycT = 0, int64_t ycells = 0) {
    using T = float;
    HE2Args<T> a;
    memset(&a, 0, sizeof(a));
    if (ycorr) {
        a.ycorr = {ycorr, (uint32_t)(ycells * 4)};
        a.ycT = ycT;
    }
    fill_geom<T>(P, P.lhe2, f, a.g);
    const uint32_t nb = (uint32_t)(P.N * (int64_t)sizeof(T));
    a.E = {(const T*)f->E, nb};
    a.H = {(const T*)f->H, nb};
    a.eps = {(const T*)f->eps, nb};
    a.E_out = {(T*)f->E_out, nb};
    a.H_out = {(T*)f->H_out, nb};
    a.S = {(T*)f->S, nb};
    a.psi_h_in = (const T*)f->psi_h;
    a.psi_h_out = (T*)f->psi_h_out;
    a.psi_e_in = (const T*)f->psi_e;
    a.psi_e_out = (T*)f->psi_e_out;
    a.coef_elems = (int)P.coef_elems;
    a.src_amp = src_amp;
    if (src_amp) {
        const int64_t N = P.N;
        const int64_t comp = src_index / N;
        const int64_t i = src_index - comp * N;
        a.src_comp = (int)comp;
        a.src_x = (int)(i / P.plane);
        a.src_p = (int)(i - (int64_t)a.src_x * P.plane);
    }
    const dim3 grid(P.lhe2.pblocks * P.lhe2.nchunk);
    const bool nu = P.d.nonuniform != 0;
    constexpr bool LC = FDTD_HE2_LDSCARRY != 0, EA = FDTD_HE2_EARLY != 0, S2 = FDTD_HE2_S2 != 0;
    // E^n neighbours from LDS (measured, per step: rows of 73 in tiles of
    // 14 rows 2.8 % faster; 64-cell z segments of rows of 160 1.1 % faster;
    // rows of 136 in tiles of 7 rows 1 % slower, where a quarter of the
    // waves hold a last row and the pass spills more)
    const bool whole = FDTD_HE2_LDSNB != 0 && (P.lhe2.tzb > 1 || P.lhe2.tyo + 3 >= 10);
#define FDTD_HE2_GO(NU_, S2_, E2_, POY_, SH_)                                                               \
    do {                                                                                                   \
        if (whole)                                                                                         \
            hipLaunchKernelGGL((he2_march<T, NU_, S2_, E2_, LC, EA, POY_, SH_, true>), grid, dim3(FDTD_HE2_HB), \
                               0, st, a);                                                                  \
        else                                                                                               \
            hipLaunchKernelGGL((he2_march<T, NU_, S2_, E2_, LC, EA, POY_, SH_, false>), grid,                 \
                               dim3(FDTD_HE2_HB), 0, st, a);                                               \
    } while (0)
    if (ycorr) {
        // the learned shell's passes (uniform grids), Poynting fused or off
        if (nu) return fail(3, "the shell correction's two-step pass needs a uniform grid");
        const bool poy = P.d.poynting == FDTD_POYNTING_FUSED;
        if (poy && last_e) FDTD_HE2_GO(false, false, true, true, true);
        else if (poy) FDTD_HE2_GO(false, false, false, true, true);
        else if (last_e) FDTD_HE2_GO(false, false, true, false, true);
        else FDTD_HE2_GO(false, false, false, false, true);
        return check_launch("fdtd_update_he2");
    }
    if (P.d.poynting == FDTD_POYNTING_FUSED) {
        if (nu && last_e) FDTD_HE2_GO(true, S2, true, true, false);
        else if (nu) FDTD_HE2_GO(true, S2, false, true, false);
        else if (last_e) FDTD_HE2_GO(false, S2, true, true, false);
        else FDTD_HE2_GO(false, S2, false, true, false);
    } else {
        // Poynting off: no S (and no S of the first half-step)
        if (nu && last_e) FDTD_HE2_GO(true, false, true, false, false);
        else if (nu) FDTD_HE2_GO(true, false, false, false, false);
        else if (last_e) FDTD_HE2_GO(false, false, true, false, false);
        else FDTD_HE2_GO(false, false, false, false, false);
    }
#undef FDTD_HE2_GO
    return check_launch("fdtd_update_he2");
}

int check_he2_fields(const Plan& P, const fdtd_fields* f) {
    if (!he2_supported(P))
        return fail(3, "fdtd_update_he2 needs FP32, fused Poynting or Poynting off, no Bloch boundaries, at most %d "
                       "CPML coefficients and no two overlapping slabs on one axis", kCoefTab);
    if (!f->E_out || f->E_out == f->E) return fail(2, "fdtd_update_he2 needs E_out != E");
    if (P.d.n_face > 0 && (!f->psi_e_out || f->psi_e_out == f->psi_e))
        return fail(2, "fdtd_update_he2 needs psi_e_out != psi_e");
    return 0;
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
    if (d.dtype != FDTD_F32 && d.dtype != FDTD_F64) return fail(1, "dtype %d unsupported (0=f32, 1=f64)", d.dtype);
    const int64_t N = d.shape[0] * d.shape[1] * d.shape[2];
    const int64_t esz = d.dtype == FDTD_F32 ? 4 : 8;
    // every per-component byte offset must fit the 32-bit offsets of the kernels
    if (N * esz > (int64_t)UINT32_MAX)
        return fail(1, "grid of %lld cells exceeds 4 GiB per field component", (long long)N);
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
        pf.face_elems = (int)(plane_of(*P, f.axis) * f.thickness);
        pf.psi_off = (int)psi_off;
        psi_off += 2 * (int64_t)pf.face_elems;
        pf.coef_off = (int)coef_off;
        coef_off += 6 * f.thickness;
        if (psi_off * esz > (int64_t)UINT32_MAX) { delete P; return fail(1, "psi block exceeds 4 GiB"); }
    }
    P->psi_elems = psi_off;
    P->coef_elems = coef_off;
    P->le = make_launch(P->plane, P->sx, kBlock, kBlock, 0, kTargetBlocks);
    if (d.poynting == FDTD_POYNTING_FUSED) {
        // LDS exchange for rows of up to FDTD_ROW_MAX cells: whole rows per workgroup
        // plus one halo row, at most FDTD_H_TARGET workgroups (as many x
        // chunks per tile column as fit); longer rows: per-cell
        // re-evaluation of the in-plane neighbours
        if (P->sz <= FDTD_ROW_MAX) {
            const int hb = P->sz <= 128 ? FDTD_HB_SMALL : 1024;
            P->lh = make_round_launch(P->plane, P->sx, hb, ((hb - P->sz) / P->sz) * P->sz, FDTD_H_TARGET);
        } else {
            P->lh = make_launch(P->plane, P->sx, kBlock, kBlock, 3, kTargetBlocks);
        }
    } else {
        P->lh = make_launch(P->plane, P->sx, kBlock, kBlock, d.poynting == FDTD_POYNTING_TWO_PASS ? 1 : 0,
                            kTargetBlocks);
    }
    // fused H+E pass: the same tile shape (whole rows plus one halo row)
    P->lhe.mode = -1;
    const bool pair_ok = d.bloch_mask == 0 || (d.dtype == FDTD_F32 && !d.nonuniform && P->coef_elems <= kCoefTab);
    const bool fused_s = d.poynting == FDTD_POYNTING_FUSED || (d.poynting == FDTD_POYNTING_OFF && !d.bloch_mask);
    if (fused_s && pair_ok && (d.bloch_mask == 0 || P->sz <= FDTD_ROW_MAX) && 3 * N * esz <= (int64_t)UINT32_MAX) {
        // one workgroup per CU at a time (1024 threads, <= 128 VGPRs); one
        // halo row (and column); the Bloch-pair pass takes whole rows only
        const int hb = P->sz <= 128 ? FDTD_HE_HB_SMALL : 1024;
        P->lhe = make_tile_launch(P->sx, P->sy, P->sz, hb, 1, 1, FDTD_HE_TARGET, 1, d.bloch_mask != 0);
    }
    // two-step pass: FP32, 1024-thread tiles with three halo rows (and,
    // with z segments, three halo columns), the coefficient block in LDS,
    // slabs of one axis apart (their psi pairs are carried between the
    // half-steps one slab per axis)
    P->lhe2.mode = -1;
    bool apart = true;
    for (int k = 0; k < d.n_face; ++k)
        for (int j = k + 1; j < d.n_face; ++j)
            if (d.face[k].axis == d.face[j].axis && P->face[k].lo < P->face[j].hi && P->face[j].lo < P->face[k].hi)
                apart = false;
    if (FDTD_HE2 && P->lhe.mode == 2 && d.bloch_mask == 0 && d.dtype == FDTD_F32 && P->coef_elems <= kCoefTab &&
        apart)
        P->lhe2 = make_tile_launch(P->sx, P->sy, P->sz, FDTD_HE2_HB, 3, 3, FDTD_HE_TARGET, 3, false);
    // Time per step of each pass: rounds of workgroups x planes a chunk
    // marches, the two-step pass's plane iteration costing FDTD_HE2_COST/10
    // of the fused pass's (measured on MI355X: 1.15 with rows of 73 cells,
    // 1.10 with rows of 128) and covering two steps.  Where the two-step
    // tiles own few of their cells (three halo rows and columns against
    // one), the cost model keeps the fused pass.
    P->he2_pref = false;
    if (P->lhe2.mode == 2) {
        const auto cost = [&](const Launch& L, int extra) {
            const int64_t blocks = (int64_t)L.pblocks * L.nchunk;
            return (double)((blocks + FDTD_HE_TARGET - 1) / FDTD_HE_TARGET) * (L.xchunk + extra);
        };
        P->he2_pref = cost(P->lhe2, 3) * (FDTD_HE2_COST / 10.0) / 2.0 < cost(P->lhe, 1);
    }
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
    const dim3 grid((P.plane + kBlock - 1) / kBlock, P.sx);
    if (P.d.dtype == FDTD_F32) {
        HArgs<float> a;
        fill_h<float>(P, P.lh, f, a);
        hipLaunchKernelGGL(s_from_os<float>, grid, dim3(kBlock), 0, st, a);
    } else {
        HArgs<double> a;
        fill_h<double>(P, P.lh, f, a);
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

}  // extern "C"

namespace {

int plane_sum_many(const Plan& P, const void* field, int comp, int axis, int64_t index, double* partials,
                   double* acc, hipStream_t st);   // adjoint.inl

// Device-to-device copy on the stream (a recorded pseudo-launch in the
// launch-recording build).
int copy_field(void* dst, const void* src, size_t bytes, hipStream_t st) {
#ifdef FDTD_RECORD_LAUNCHES
    fdtd_record_copy(dst, src, bytes);
    (void)st;
#else
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(5, "hipMemcpyAsync of %zu bytes failed", bytes);
#endif
    return 0;
}

// n steps of the caller loop.  home: E and psi_e end in f->E and f->psi_e
// (pass counts kept even); otherwise they may end in f->E_out / f->psi_e_out
// and *moved says where (bit 0: E, bit 1: psi_e).  obs: the observations of
// fdtd_step_n_observed after every step (one step per pass then).
// flags (fdtd_step_n_chain): FDTD_CHAIN_H_PENDING runs the previous
// call's last H half-step first, FDTD_CHAIN_DEFER_H leaves this call's.
int step_n_impl(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index, const void* src_wave,
                void* stream, bool home, int32_t* moved, const fdtd_observe* obs = nullptr, int32_t flags = 0) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (moved) *moved = 0;
    if (n < 0) return fail(3, "n = %d < 0", n);
    if (flags & ~(FDTD_CHAIN_H_PENDING | FDTD_CHAIN_DEFER_H)) return fail(3, "unknown chain flags 0x%x", flags);
    const bool pend_in = (flags & FDTD_CHAIN_H_PENDING) != 0;
    const bool defer = (flags & FDTD_CHAIN_DEFER_H) != 0 && n > 0;
    if ((pend_in || defer) && (home || obs)) return fail(3, "chain flags need fdtd_step_n_chain");
    if (src_wave && (src_index < 0 || src_index >= 3 * P.N))
        return fail(3, "src_index %lld out of [0, %lld)", (long long)src_index, (long long)(3 * P.N));
    if (int rc = check_fields(P, f, true)) return rc;
    if (obs && obs->flux_acc) {
        if (P.d.poynting == FDTD_POYNTING_OFF || !f->S) return fail(3, "the flux needs S (Poynting on)");
        if (obs->flux_comp < 0 || obs->flux_comp > 2 || obs->flux_axis < 0 || obs->flux_axis > 2)
            return fail(3, "flux component %d / axis %d out of [0, 3)", obs->flux_comp, obs->flux_axis);
        const int64_t s = obs->flux_axis == 0 ? P.sx : obs->flux_axis == 1 ? P.sy : P.sz;
        if (obs->flux_index < 0 || obs->flux_index >= s)
            return fail(3, "flux plane %lld out of [0, %lld)", (long long)obs->flux_index, (long long)s);
    }
    hipStream_t st = (hipStream_t)stream;
    fdtd_fields cur = *f;
    const size_t esz = P.d.dtype == FDTD_F32 ? 4 : 8;
    const bool swap = (f->H_out && f->H_out != f->H);
    const auto amp_of = [&](int k) -> const void* {
        return src_wave ? (const char*)src_wave + (size_t)k * esz : nullptr;
    };
    // observations of step k: E once its E half-step (and source) is done,
    // the flux and H once its H half-step (and S) is
    const size_t fbytes = (size_t)(3 * P.N) * esz;
    const auto seen_e = [&](int k) -> int {
        return (obs && obs->E_keep) ? copy_field(obs->E_keep[k], cur.E, fbytes, st) : 0;
    };
    const auto seen_h = [&](int k) -> int {
        if (obs && obs->flux_acc) {
            const int rc = obs->flux_partials
                               ? plane_sum_many(P, cur.S, obs->flux_comp, obs->flux_axis, obs->flux_index,
                                                obs->flux_partials, obs->flux_acc, st)
                               : fdtd_plane_sum(plan, cur.S, obs->flux_comp, obs->flux_axis, obs->flux_index,
                                                obs->flux_acc, stream);
            if (rc) return rc;
        }
        return (obs && obs->H_keep) ? copy_field(obs->H_keep[k], cur.H, fbytes, st) : 0;
    };
    const auto half_e = [&](int k) {
        return (P.d.dtype == FDTD_F32) ? launch_e<float>(P, &cur, st, (const float*)amp_of(k), src_index)
                                       : launch_e<double>(P, &cur, st, (const double*)amp_of(k), src_index);
    };
    const auto half_h = [&]() {
        int rc = (P.d.dtype == FDTD_F32) ? launch_h<float>(P, &cur, st) : launch_h<double>(P, &cur, st);
        if (rc == 0 && P.d.poynting == FDTD_POYNTING_TWO_PASS) rc = fdtd_update_s(plan, &cur, stream);
        if (swap) {
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
            if (P.d.bloch_mask && cur.G_out && cur.G_out != cur.G) {
                std::swap(cur.G, cur.G_out);
                std::swap(cur.psi_g, cur.psi_g_out);
            }
        }
        return rc;
    };
    // Half-steps E_0 H_0 E_1 H_1 ... E_{n-1} H_{n-1}.  Fused: E_0, then the
    // m = n-1 pairs (H_{k-1}, E_k), then H_{n-1}.  A pair runs unfused, in
    // one he_march, or two pairs in one he2_march.  Each he_march moves E,
    // and each he2_march E and psi_e, to the other buffer and H once.  home:
    // both counts are kept even (E and psi_e end in f->E and f->psi_e, and
    // H swaps n times in all, as unfused): 2 * (m / 4) he2_march passes,
    // then of the rest an even number of he_march and at most one unfused.
    // Otherwise m / 2 he2_march passes and at most one he_march.  With a
    // pending H_{-1} the pairs start at (H_{-1}, E_0): m = n; with a
    // deferred H_{n-1} the schedule ends after the last pair.
    const bool pair_bufs = !P.d.bloch_mask ||
                           (f->F_out && f->F_out != f->F && f->G_out && f->G_out != f->G &&
                            (P.d.n_face == 0 || (f->psi_g_out && f->psi_g_out != f->psi_g)));
    const bool fuse = n >= (pend_in ? 1 : 2) && he_buffers(P, f) && he_supported(P) && pair_bufs;
    const bool fuse2 = fuse && !obs && he2_supported(P) && P.he2_pref &&
                       (P.d.n_face == 0 || (f->psi_e_out && f->psi_e_out != f->psi_e));
    int k = 0;
    if (fuse) {
        if (!pend_in) {
            if (int rc = half_e(0)) return rc;
            if (int rc = seen_e(0)) return rc;
            k = 1;
        }
        const int m = pend_in ? n : n - 1;
        const int n2 = fuse2 ? (home ? 2 * (m / 4) : m / 2) : 0;
        // shortest schedule with an odd pair count: the last pair and the
        // final H half-step run as one pass (he2_march without its last E)
        const bool heh = fuse2 && !home && !defer && (m - 2 * n2) == 1;
        const int r = m - 2 * n2 - (heh ? 1 : 0);
        if (home && r % 2 == 1) {
            if (int rc = half_h()) return rc;
            if (int rc = seen_h(0)) return rc;
            if (int rc = half_e(1)) return rc;
            if (int rc = seen_e(1)) return rc;
            k = 2;
        }
        for (int j = 0; j < (home ? r / 2 * 2 : r); ++j, ++k) {
            int rc = (P.d.dtype == FDTD_F32) ? launch_he<float>(P, &cur, st, (const float*)amp_of(k), src_index)
                                             : launch_he<double>(P, &cur, st, (const double*)amp_of(k), src_index);
            if (rc) return rc;
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
            std::swap(cur.E, cur.E_out);
            if (P.d.bloch_mask) {
                std::swap(cur.F, cur.F_out);
                std::swap(cur.G, cur.G_out);
                std::swap(cur.psi_g, cur.psi_g_out);
            }
            if (int rc = seen_h(k - 1)) return rc;
            if (int rc = seen_e(k)) return rc;
        }
        for (int j = 0; j < n2 + (heh ? 1 : 0); ++j, k += 2) {
            const bool last_e = !(heh && j == n2);
            if (int rc = launch_he2(P, &cur, st, (const float*)amp_of(k), src_index, last_e)) return rc;
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
            std::swap(cur.E, cur.E_out);
            std::swap(cur.psi_e, cur.psi_e_out);
        }
        if (!heh && !defer) {
            if (int rc = half_h()) return rc;
            if (int rc = seen_h(n - 1)) return rc;
        }
    } else {
        if (pend_in && n > 0) {
            if (int rc = half_h()) return rc;
        }
        for (k = 0; k < n; ++k) {
            if (int rc = half_e(k)) return rc;
            if (int rc = seen_e(k)) return rc;
            if (defer && k == n - 1) break;
            if (int rc = half_h()) return rc;
            if (int rc = seen_h(k)) return rc;
        }
    }
    // a two-step pass swaps H once for two H half-steps, so with an odd
    // number of them the newest H is not where n's parity puts it
    if (moved)
        *moved = (cur.E != f->E ? 1 : 0) | (cur.psi_e != f->psi_e ? 2 : 0) | (cur.H != f->H ? 4 : 0) |
                 ((defer || (pend_in && n == 0)) ? FDTD_MOVED_H_PENDING : 0);
    return 0;
}

}  // namespace

extern "C" {

int fdtd_step_n(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index, const void* src_wave,
                void* stream) {
    return step_n_impl(plan, f, n, src_index, src_wave, stream, true, nullptr);
}

int fdtd_step_n_moved(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index, const void* src_wave,
                      void* stream, int32_t* moved) {
    if (!moved) return fail(2, "moved is NULL");
    return step_n_impl(plan, f, n, src_index, src_wave, stream, false, moved);
}

int fdtd_step_n_chain(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index, const void* src_wave,
                      int32_t flags, void* stream, int32_t* moved) {
    if (!moved) return fail(2, "moved is NULL");
    return step_n_impl(plan, f, n, src_index, src_wave, stream, false, moved, nullptr, flags);
}

int fdtd_step_n_observed(const void* plan, const fdtd_fields* f, int32_t n, int64_t src_index,
                         const void* src_wave, const fdtd_observe* obs, void* stream, int32_t* moved) {
    if (!moved) return fail(2, "moved is NULL");
    if (obs)
        for (int k = 0; k < n; ++k)
            if ((obs->E_keep && !obs->E_keep[k]) || (obs->H_keep && !obs->H_keep[k]))
                return fail(2, "E_keep / H_keep entry %d is NULL", k);
    return step_n_impl(plan, f, n, src_index, src_wave, stream, false, moved, obs);
}

int fdtd_update_he(const void* plan, const fdtd_fields* f, int64_t src_index, const void* src_amp,
                   void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_fields(P, f, true)) return rc;
    if (int rc = check_he_fields(P, f)) return rc;
    if (src_amp && (src_index < 0 || src_index >= 3 * P.N))
        return fail(3, "src_index %lld out of [0, %lld)", (long long)src_index, (long long)(3 * P.N));
    hipStream_t st = (hipStream_t)stream;
    if (P.d.bloch_mask) return launch_heb(P, f, st, (const float*)src_amp, src_index);
    if (P.d.dtype == FDTD_F32) return launch_he<float>(P, f, st, (const float*)src_amp, src_index);
    return launch_he<double>(P, f, st, (const double*)src_amp, src_index);
}

int fdtd_update_he_supported(const void* plan) { return plan && he_supported(*(const Plan*)plan) ? 1 : 0; }

int fdtd_update_he2(const void* plan, const fdtd_fields* f, int64_t src_index, const void* src_amp,
                    void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_fields(P, f, true)) return rc;
    if (int rc = check_he2_fields(P, f)) return rc;
    if (src_amp && (src_index < 0 || src_index >= 3 * P.N))
        return fail(3, "src_index %lld out of [0, %lld)", (long long)src_index, (long long)(3 * P.N));
    return launch_he2(P, f, (hipStream_t)stream, (const float*)src_amp, src_index);
}

int fdtd_update_he2_supported(const void* plan) { return plan && he2_supported(*(const Plan*)plan) ? 1 : 0; }

}  // extern "C"

namespace {

std::string launch_text(const Launch& L) {
    char b[200];
    if (L.tzb > 1)
        snprintf(b, sizeof(b),
                 "%d workgroups of %d threads (%d tile columns of %d rows x %d-cell z segments x %d x-chunks of "
                 "%d planes)",
                 L.pblocks * L.nchunk, L.block, L.pblocks, L.tyo, L.tzo, L.nchunk, L.xchunk);
    else
        snprintf(b, sizeof(b),
                 "%d workgroups of %d threads (%d tile columns of %d plane cells x %d x-chunks of %d planes)",
                 L.pblocks * L.nchunk, L.block, L.pblocks, L.tile, L.nchunk, L.xchunk);
    return b;
}

// The specialisation behind one update of the plan: what the reference's
// rendered template source held (simulation.py:202-230: the E/H/S/F/G
// templates rendered with pmls, do_poynting, bloch and uniform_dx), as the
// kernel, its template flags, launch shape and face list chosen here.
std::string describe(const Plan& P, int which) {
    const fdtd_desc& d = P.d;
    const char* ty = d.dtype == FDTD_F32 ? "float" : "double";
    const bool pair = d.bloch_mask != 0;
    const bool seam = pair && d.bloch_mode == FDTD_BLOCH_PHASE;
    char b[512];
    std::string out;
    const char* what[] = {"update_E", "update_H", "update_S", "update_F", "update_G", "step_n"};
    snprintf(b, sizeof(b), "/* %s: libfdtd_hip.so, gfx950; grid %d x %d x %d (%lld cells), %s, dt = %.17g */\n",
             what[which], P.sx, P.sy, P.sz, (long long)P.N, ty, d.dt);
    out += b;
    if (d.nonuniform) out += "/* dx: per-index 1/dx_e (H step) and 1/dx_h (E step) arrays */\n";
    else {
        snprintf(b, sizeof(b), "/* dx: uniform, 1/dx = %.17g */\n", d.inv_dx);
        out += b;
    }
    for (int k = 0; k < d.n_face; ++k) {
        const fdtd_face& f = d.face[k];
        snprintf(b, sizeof(b), "/* pml %d: axis %c, polarity %c, thickness %d, slab %d..%d, psi offset %d */\n", k,
                 "xyz"[f.axis], f.polarity ? 'p' : 'n', f.thickness, P.face[k].lo, P.face[k].hi - 1,
                 P.face[k].psi_off);
        out += b;
    }
    for (int r = 0; r < 3; ++r)
        if ((d.bloch_mask >> r) & 1) {
            snprintf(b, sizeof(b), "/* bloch: axis %c, real %.17g, imag %.17g (%s) */\n", "xyz"[r], d.bloch_re[r],
                     d.bloch_im[r], seam ? "phase applied at the seam" : "decoupled, as the reference renders it");
            out += b;
        }
    const char* pm[] = {"off", "fused", "two_pass"};
    switch (which) {
        case 0:
        case 3:
            snprintf(b, sizeof(b), "e_march<%s, PAIR=%d, SEAM=%d>: %s%s\n", ty, (int)pair, (int)seam,
                     launch_text(P.le).c_str(), which == 3 ? "; F is advanced by update_E's paired launch" : "");
            out += b;
            break;
        case 1:
        case 4:
            snprintf(b, sizeof(b), "h_march<%s, PAIR=%d, SEAM=%d, MODE=%d>: Poynting %s; %s%s\n", ty, (int)pair,
                     (int)seam, P.lh.mode, pm[d.poynting], launch_text(P.lh).c_str(),
                     which == 4 ? "; G is advanced by update_H's paired launch" : "");
            out += b;
            break;
        case 2:
            if (d.poynting == FDTD_POYNTING_TWO_PASS)
                snprintf(b, sizeof(b), "s_from_os<%s>: S from the 6N oS products, one thread per cell\n", ty);
            else
                snprintf(b, sizeof(b), "S is computed by h_march (fused Poynting); update_S enqueues nothing\n");
            out += b;
            break;
        default:
            if (he_supported(P)) {
                out += "he_march<" + std::string(ty) + ">: " +
                       (d.poynting == FDTD_POYNTING_OFF ? "H then E" : "H+S then E") + " in one pass; " +
                       launch_text(P.lhe) + "\n";
                if (he2_supported(P)) {
                    out += "he2_march<float>: two steps per pass; " + launch_text(P.lhe2) +
                           (P.he2_pref ? " (chosen by the time model)\n" : " (not chosen: the one-step pass is faster)\n");
                }
            } else {
                out += "step_n runs update_E and update_H per step (no fused pass for this plan)\n";
            }
    }
    return out;
}

}  // namespace

extern "C" {

int64_t fdtd_plan_describe(const void* plan, int32_t which, char* buf, int64_t size) {
    if (!plan) return fail(1, "plan is NULL"), -1;
    if (which < 0 || which > 5) return fail(3, "which = %d not in 0..5", which), -1;
    const std::string s = describe(*(const Plan*)plan, which);
    if (buf && size > 0) {
        const int64_t n = (int64_t)s.size() < size - 1 ? (int64_t)s.size() : size - 1;
        memcpy(buf, s.data(), (size_t)n);
        buf[n] = 0;
    }
    return (int64_t)s.size();
}

int fdtd_update_he2_preferred(const void* plan) {
    return plan && he2_supported(*(const Plan*)plan) && ((const Plan*)plan)->he2_pref ? 1 : 0;
}

}  // extern "C"

#include "adjoint.inl"

#include "shell_net.inl"
