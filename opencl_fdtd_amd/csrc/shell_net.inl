// The learned absorbing shell (part of the fdtd_hip.hip translation unit).
//
// A thin shell of T cells on every face runs the CPML-T update; after each E
// update (and source) a small convolutional net adds a correction to E on the
// shell cells:
//   x = (Ex, Ey, Ez, Hx, Hy, Hz)                         6 input channels
//   h_k(j) = sum_{c,t} W1[k][c][t] x_c(j + t)            16 hidden channels
//   y_a(i) = sum_{k,t} W2[d(i)][a][k][t] h_k(i + t)      3 outputs
//   E_a(i) += y_a(i)  for i in the shell
// with d(i) the depth of i: its index along its box's thin axis (a layer-2
// head per depth).
// t runs over the 27 offsets of a 3x3x3 stencil (periodic wrap, as the step).
// There is no activation: the correction must be linear in the fields, as
// Maxwell's equations are.  Weights and activations are BF16 (inputs rounded
// to BF16), products are exact in FP32 and accumulated in FP32.
//
// The shell is split into 6 boxes, one per face (x faces own the edges and
// corners they touch, then y faces), each with its own weights.  Hidden values
// live per box on the box grown by one cell, in a caller-provided FP32
// scratch [cell][16].

#ifndef FDTD_SHELL_TILE_A1
#define FDTD_SHELL_TILE_A1 8         // shell_mfma tile: output rows along a1 (even; 8 or 10)
#endif
#ifndef FDTD_SHELL_WAVES
#define FDTD_SHELL_WAVES 4           // shell_mfma: minimum waves per SIMD (4 workgroups of 4 waves per CU)
#endif

namespace {

constexpr int kHid = 16;
constexpr int kIn = 6;
constexpr int kOut = 3;
constexpr int kTaps = 27;
constexpr int kDepth = 4;                   // layer-2 heads per box, one per shell depth (T <= 4)
constexpr int kW1 = kHid * kIn * kTaps;    // per box
constexpr int kW2 = kOut * kHid * kTaps;   // per head

// Along an axis the box spans completely (full), its grown regions are the
// same periodic range (local coordinate = global coordinate, wrapping);
// along the others they grow by 1 (hidden) or 2 (input adjoint) cells on
// each side.
struct ShellBox {
    int o0[3], o1[3];    // output region [o0, o1) per axis
    int full[3];         // 1: o0 = 0 and o1 = the grid size
    int s[3];            // grid size per axis
    int e[3];            // hidden (grown) extents
    int hoff;            // offset of the box's hidden block, in cells
    int hcells;          // e[0]*e[1]*e[2]
};

// global coordinate of hidden-local l, and hidden-local of global g (-1 if outside)
__device__ __forceinline__ int hid_global(const ShellBox& B, int q, int l) {
    return B.full[q] ? l : B.o0[q] - 1 + l;
}
__device__ __forceinline__ int hid_local(const ShellBox& B, int q, int g) {
    if (B.full[q]) return g < 0 ? g + B.s[q] : (g >= B.s[q] ? g - B.s[q] : g);
    const int l = g - (B.o0[q] - 1);
    return (l < 0 || l >= B.e[q]) ? -1 : l;
}
// output-local of global g (-1 if outside)
__device__ __forceinline__ int out_local(const ShellBox& B, int q, int g) {
    if (B.full[q]) return g < 0 ? g + B.s[q] : (g >= B.s[q] ? g - B.s[q] : g);
    const int l = g - B.o0[q];
    return (l < 0 || l >= B.o1[q] - B.o0[q]) ? -1 : l;
}
// extent and origin of the input-adjoint region (grown by 2)
__device__ __forceinline__ int in_extent(const ShellBox& B, int q) {
    return B.full[q] ? B.s[q] : B.o1[q] - B.o0[q] + 4;
}
__device__ __forceinline__ int in_global(const ShellBox& B, int q, int l) {
    return B.full[q] ? l : B.o0[q] - 2 + l;
}

template <class T>
struct ShellArgs {
    int sx, sy, sz, nbox;
    ShellBox box[6];
    const uint16_t* w;   // BF16: W1 for all boxes, then W2 for all boxes
    const T* E;          // input E (the correction reads E before it is applied)
    const T* H;
    T* Eout;             // E to correct (may equal E: layer 2 reads only hidden)
    float* hid;          // hidden scratch
    float* hbar;         // hidden adjoint scratch
    const T* Ebar;       // adjoint of corrected E (backward)
    T* Ebar_out;         // adjoint of E before the correction (accumulated)
    T* Hbar_out;         // adjoint of H (accumulated)
    float* wgrad_part;   // per-workgroup weight-gradient partials
    int part_box;        // box handled by a per-box launch
    int nparts;
};

__device__ __forceinline__ float bf16_to_f(uint16_t v) {
    return __uint_as_float((uint32_t)v << 16);
}
// round-to-nearest-even to BF16, returned as the FP32 value it represents
__device__ __forceinline__ float bf16_round(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return f;   // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

__device__ __forceinline__ int wrapi(int c, int s) { return c < 0 ? c + s : (c >= s ? c - s : c); }

template <class T>
__device__ __forceinline__ size_t gidx(const ShellArgs<T>& a, int x, int y, int z) {
    return ((size_t)wrapi(x, a.sx) * a.sy + wrapi(y, a.sy)) * a.sz + wrapi(z, a.sz);
}

// global linear cell of the grown box -> (box, local coords)
template <class T>
__device__ __forceinline__ bool locate(const ShellArgs<T>& a, int cell, int* b, int* lx, int* ly, int* lz) {
    for (int k = 0; k < a.nbox; ++k) {
        const ShellBox& B = a.box[k];
        if (cell >= B.hoff && cell < B.hoff + B.hcells) {
            const int l = cell - B.hoff;
            *b = k;
            *lz = l % B.e[2];
            *ly = (l / B.e[2]) % B.e[1];
            *lx = l / (B.e[2] * B.e[1]);
            return true;
        }
    }
    return false;
}

// layer 1 over every grown-box cell
template <class T>
__global__ __launch_bounds__(256) void shell_l1(ShellArgs<T> a, int total) {
    const int cell = blockIdx.x * 256 + threadIdx.x;
    int b, lx, ly, lz;
    if (cell >= total || !locate(a, cell, &b, &lx, &ly, &lz)) return;
    const ShellBox& B = a.box[b];
    const int x = hid_global(B, 0, lx), y = hid_global(B, 1, ly), z = hid_global(B, 2, lz);
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const uint16_t* w1 = a.w + (size_t)b * kW1;
    float acc[kHid];
#pragma unroll
    for (int k = 0; k < kHid; ++k) acc[k] = 0.f;
    for (int t = 0; t < kTaps; ++t) {
        const size_t gi = gidx(a, x + t / 9 - 1, y + (t / 3) % 3 - 1, z + t % 3 - 1);
        float xv[kIn];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xv[c] = bf16_round((float)a.E[c * N + gi]);
            xv[3 + c] = bf16_round((float)a.H[c * N + gi]);
        }
#pragma unroll
        for (int k = 0; k < kHid; ++k)
#pragma unroll
            for (int c = 0; c < kIn; ++c) acc[k] += bf16_to_f(w1[(k * kIn + c) * kTaps + t]) * xv[c];
    }
    float* h = a.hid + (size_t)cell * kHid;
#pragma unroll
    for (int k = 0; k < kHid; ++k) h[k] = bf16_round(acc[k]);
}

// ---- backward (training): adjoint of E_out = E + mask * y(E, H) ----------

// hbar_k(j) = sum_{o,t} W2[o][k][t] ybar_o(j - t), ybar = Ebar on the box's
// output cells (0 elsewhere); j over the grown box
template <class T>
__global__ __launch_bounds__(256) void shell_l2_t(ShellArgs<T> a, int total) {
    const int cell = blockIdx.x * 256 + threadIdx.x;
    int b, lx, ly, lz;
    if (cell >= total || !locate(a, cell, &b, &lx, &ly, &lz)) return;
    const ShellBox& B = a.box[b];
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const uint16_t* w2b = a.w + (size_t)a.nbox * kW1 + (size_t)b * kDepth * kW2;
    const int r = b / 2;                       // the box's thin axis: depth = output-local index along it
    float acc[kHid];
#pragma unroll
    for (int k = 0; k < kHid; ++k) acc[k] = 0.f;
    const int gx = hid_global(B, 0, lx), gy = hid_global(B, 1, ly), gz = hid_global(B, 2, lz);
    for (int t = 0; t < kTaps; ++t) {
        // output cell i = j - t
        const int ox = out_local(B, 0, gx - (t / 9 - 1));
        const int oy = out_local(B, 1, gy - ((t / 3) % 3 - 1));
        const int oz = out_local(B, 2, gz - (t % 3 - 1));
        if (ox < 0 || oy < 0 || oz < 0) continue;
        const uint16_t* w2 = w2b + (size_t)(r == 0 ? ox : (r == 1 ? oy : oz)) * kW2;
        const size_t gi = gidx(a, B.o0[0] + ox, B.o0[1] + oy, B.o0[2] + oz);
        float yb[kOut];
#pragma unroll
        for (int o = 0; o < kOut; ++o) yb[o] = (float)a.Ebar[o * N + gi];
#pragma unroll
        for (int k = 0; k < kHid; ++k)
#pragma unroll
            for (int o = 0; o < kOut; ++o) acc[k] += bf16_to_f(w2[(o * kHid + k) * kTaps + t]) * yb[o];
    }
    float* hb = a.hbar + (size_t)cell * kHid;
#pragma unroll
    for (int k = 0; k < kHid; ++k) hb[k] = acc[k];
}

// xbar_c(i) = sum_{k,t} W1[k][c][t] hbar_k(i - t) for one box (part_box);
// i over the box grown by two cells; accumulated into Ebar_out / Hbar_out.
template <class T>
__global__ __launch_bounds__(256) void shell_l1_t(ShellArgs<T> a) {
    const ShellBox& B = a.box[a.part_box];
    const int g0 = in_extent(B, 0), g1 = in_extent(B, 1), g2 = in_extent(B, 2);
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= g0 * g1 * g2) return;
    const int lz = cell % g2, ly = (cell / g2) % g1, lx = cell / (g2 * g1);   // grown-by-2 coords
    const int gx = in_global(B, 0, lx), gy = in_global(B, 1, ly), gz = in_global(B, 2, lz);
    const uint16_t* w1 = a.w + (size_t)a.part_box * kW1;
    float acc[kIn];
#pragma unroll
    for (int c = 0; c < kIn; ++c) acc[c] = 0.f;
    for (int t = 0; t < kTaps; ++t) {
        // hidden cell j = i - t
        const int hx = hid_local(B, 0, gx - (t / 9 - 1));
        const int hy = hid_local(B, 1, gy - ((t / 3) % 3 - 1));
        const int hz = hid_local(B, 2, gz - (t % 3 - 1));
        if (hx < 0 || hy < 0 || hz < 0) continue;
        const float* hb = a.hbar + ((size_t)B.hoff + ((size_t)hx * B.e[1] + hy) * B.e[2] + hz) * kHid;
#pragma unroll
        for (int k = 0; k < kHid; ++k) {
            const float v = hb[k];
#pragma unroll
            for (int c = 0; c < kIn; ++c) acc[c] += bf16_to_f(w1[(k * kIn + c) * kTaps + t]) * v;
        }
    }
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const size_t gi = gidx(a, gx, gy, gz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a.Ebar_out[c * N + gi] += (T)acc[c];
        a.Hbar_out[c * N + gi] += (T)acc[3 + c];
    }
}

// Weight-gradient partials of one box (part_box), GEMM-shaped:
//   W1bar[k][(c,t)] = sum_j hbar_k(j) x_c(j + t)   over the grown box
//   W2bar[o][(k,t)] = sum_i ybar_o(i) h_k(i + t)   over the output box
// Workgroup g takes a contiguous slice of the cells; chunks of the slice are
// staged in LDS (adjoint rows and im2col rows) and each thread accumulates a
// fixed subset of the weights.  partials[g][w], summed in a fixed order later.
constexpr int kChunk1 = 32;
constexpr int kChunk2 = 16;

template <class T>
__global__ __launch_bounds__(256) void shell_wgrad1(ShellArgs<T> a) {
    __shared__ float sh[kChunk1][kHid];
    __shared__ float sx[kChunk1][kIn * kTaps];
    const ShellBox& B = a.box[a.part_box];
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const int per = (B.hcells + gridDim.x - 1) / gridDim.x;
    const int c0 = blockIdx.x * per, c1 = min(c0 + per, B.hcells);
    constexpr int NW = kHid * kIn * kTaps;   // 2592
    constexpr int PER = (NW + 255) / 256;
    float acc[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) acc[m] = 0.f;
    for (int base = c0; base < c1; base += kChunk1) {
        const int n = min(kChunk1, c1 - base);
        for (int e = threadIdx.x; e < kChunk1 * kHid; e += 256) {
            const int l = e / kHid, k = e % kHid;
            sh[l][k] = l < n ? a.hbar[((size_t)B.hoff + base + l) * kHid + k] : 0.f;
        }
        for (int e = threadIdx.x; e < kChunk1 * kIn * kTaps; e += 256) {
            const int l = e / (kIn * kTaps), ct = e % (kIn * kTaps), c = ct / kTaps, t = ct % kTaps;
            float v = 0.f;
            if (l < n) {
                const int cl = base + l;
                const int lz = cl % B.e[2], ly = (cl / B.e[2]) % B.e[1], lx = cl / (B.e[2] * B.e[1]);
                const size_t gi = gidx(a, hid_global(B, 0, lx) + t / 9 - 1, hid_global(B, 1, ly) + (t / 3) % 3 - 1,
                                       hid_global(B, 2, lz) + t % 3 - 1);
                v = bf16_round((float)(c < 3 ? a.E[c * N + gi] : a.H[(c - 3) * N + gi]));
            }
            sx[l][ct] = v;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int wi = threadIdx.x + 256 * m;
            if (wi < NW) {
                const int k = wi / (kIn * kTaps), ct = wi % (kIn * kTaps);
                float s_ = 0.f;
                for (int l = 0; l < kChunk1; ++l) s_ += sh[l][k] * sx[l][ct];
                acc[m] += s_;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int wi = threadIdx.x + 256 * m;
        if (wi < NW) a.wgrad_part[(size_t)blockIdx.x * NW + wi] = acc[m];
    }
}

// (layer 2: all kDepth heads, each summing only the cells at its depth)
template <class T>
__global__ __launch_bounds__(256) void shell_wgrad2(ShellArgs<T> a) {
    __shared__ float sy[kChunk2][kOut];
    __shared__ int sd[kChunk2];
    __shared__ float shc[kChunk2][kHid * kTaps];
    const ShellBox& B = a.box[a.part_box];
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    const int nout = (B.o1[0] - B.o0[0]) * ny * nz;
    const int per = (nout + gridDim.x - 1) / gridDim.x;
    const int c0 = blockIdx.x * per, c1 = min(c0 + per, nout);
    constexpr int NW = kDepth * kOut * kHid * kTaps;  // 4 x 1296
    constexpr int PER = (NW + 255) / 256;
    const int r = a.part_box / 2;
    float acc[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) acc[m] = 0.f;
    for (int base = c0; base < c1; base += kChunk2) {
        const int n = min(kChunk2, c1 - base);
        for (int e = threadIdx.x; e < kChunk2 * kOut; e += 256) {
            const int l = e / kOut, o = e % kOut;
            float v = 0.f;
            int dep = -1;
            if (l < n) {
                const int cl = base + l;
                const int lz = cl % nz, ly = (cl / nz) % ny, lx = cl / (nz * ny);
                v = (float)a.Ebar[o * N + gidx(a, B.o0[0] + lx, B.o0[1] + ly, B.o0[2] + lz)];
                dep = r == 0 ? lx : (r == 1 ? ly : lz);
            }
            sy[l][o] = v;
            if (o == 0) sd[l] = dep;
        }
        for (int e = threadIdx.x; e < kChunk2 * kHid * kTaps; e += 256) {
            const int l = e / (kHid * kTaps), kt = e % (kHid * kTaps), k = kt / kTaps, t = kt % kTaps;
            float v = 0.f;
            if (l < n) {
                const int cl = base + l;
                const int lz = cl % nz, ly = (cl / nz) % ny, lx = cl / (nz * ny);
                const int hx = hid_local(B, 0, B.o0[0] + lx + t / 9 - 1);
                const int hy = hid_local(B, 1, B.o0[1] + ly + (t / 3) % 3 - 1);
                const int hz = hid_local(B, 2, B.o0[2] + lz + t % 3 - 1);
                v = a.hid[((size_t)B.hoff + ((size_t)hx * B.e[1] + hy) * B.e[2] + hz) * kHid + k];
            }
            shc[l][kt] = v;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int wi = threadIdx.x + 256 * m;
            if (wi < NW) {
                const int dep = wi / (kOut * kHid * kTaps), w = wi % (kOut * kHid * kTaps);
                const int o = w / (kHid * kTaps), kt = w % (kHid * kTaps);
                float s_ = 0.f;
                for (int l = 0; l < kChunk2; ++l) s_ += sd[l] == dep ? sy[l][o] * shc[l][kt] : 0.f;
                acc[m] += s_;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int wi = threadIdx.x + 256 * m;
        if (wi < NW) a.wgrad_part[(size_t)blockIdx.x * NW + wi] = acc[m];
    }
}

// wgrad[b][w] += sum_g partials[g][w], fixed order
__global__ __launch_bounds__(256) void shell_wgrad_sum(const float* part, int groups, int nw, float* out) {
    const int wi = blockIdx.x * 256 + threadIdx.x;
    if (wi >= nw) return;
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += (double)part[(size_t)g * nw + wi];
    out[wi] += (float)s;
}

// ---- inference with MFMA (v_mfma_f32_16x16x32_bf16) ----------------------
//
// One workgroup (4 waves) per tile of a box.  A tile has three axes: a0 the
// box's thin axis (T cells), a1 (8 cells) and a2 (14 cells), a2 innermost in
// LDS; a2 is z for the x and y faces and y for the z faces (a1 is the
// remaining axis).  The tile's input grown by 2 (T+4 x 12 x 18 cells) is
// staged in LDS as BF16 [cell][8] (6 channels + 2 zero pads), its hidden
// layer grown by 1 (T+2 x 10 x 16) as BF16 in two planes [cell][8 channels]
// over the same LDS words.
//
// Each layer is a GEMM with the weights as the A operand (M = the layer's
// output channels) and 16 cells of one a2 row as B (N; a hidden row is
// exactly 16 cells, an output row 14 and two discarded), K running over
// (tap, input channel).  A fragment (K = 32) covers one tap pair (d0, d1)
// along a0, a1 with its three taps along a2 in the lane groups:
//   layer 1: 9 fragments, group g < 3 = tap d2 = g x 8 input channels
//            (6 used), group 3 zero;
//   layer 2: 18 fragments, two per pair: (d2 = 0 channels 0-7, 8-15,
//            d2 = 1 channels 0-7, 8-15) and (d2 = 2 channels 0-7, 8-15,
//            zero, zero).
// The fragment of pair (d0, d1) for row r reads the same cells as pair
// (d0, d1 + 1) for row r - 1, so a wave walking a run of rows along a1 loads
// each B fragment once and feeds it to up to 3 rows' MFMAs.
// MFMA operands (16x16x32): lane l holds A row l%16 and B column l%16 at
// k = 8*(l/16) + [0..7]; D column l%16 at rows 4*(l/16) + [0..3], i.e. for
// cell l%16 the output channels 4*(l/16) + [0..3].  The correction goes to a
// scratch in box order [3][shell cells] and is added to E later (by the next
// E update, or by shell_add), since neighbouring tiles read E around their
// own cells.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxT = 4;
constexpr int kTileA1 = FDTD_SHELL_TILE_A1, kTileA2 = 14;   // output cells along a1, a2
constexpr int kRow = 16;                                  // MFMA columns = hidden a2 extent
constexpr int kInA0 = (kTileA1 + 4) * (kTileA2 + 4) + 1;  // input a0 stride (padded: no bank conflicts)
constexpr int kInMax = (kMaxT + 4) * kInA0;
constexpr int kHidRows = (kMaxT + 2) * (kTileA1 + 2);     // hidden a2 rows (of 16 cells)
constexpr int kHidPlane = kHidRows * kRow;                // cells per hidden plane
constexpr int kTileLds = kInMax > 2 * kHidPlane ? kInMax : 2 * kHidPlane;   // uint4 words
static_assert(kTileA2 + 2 == kRow, "hidden rows must be one MFMA wide");
static_assert(kDepth >= kMaxT, "one layer-2 head per shell depth");

struct ShellTiles {
    int perm[6][3];    // grid axis of tile axes a0, a1, a2
    int te[6][3];      // output extents along a0, a1, a2
    int nt[6][3];      // tiles along a0, a1, a2
    int tstart[7];     // prefix of tiles per box
    int ostart[7];     // prefix of output cells per box
};

// BF16 conversions with round-to-nearest-even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    bf16x2 r;
    r[0] = (__bf16)lo;
    r[1] = (__bf16)hi;
    return __builtin_bit_cast(uint32_t, r);
}

// grid-axis displacement (0..2) of tap t along grid axis q, and the tap of
// tile-axis displacements (d0, d1, d2) for a tile with axis permutation p
__host__ __device__ __forceinline__ int tap_of(const int* p, int d0, int d1, int d2) {
    int d[3];
    d[p[0]] = d0; d[p[1]] = d1; d[p[2]] = d2;
    return (d[0] * 3 + d[1]) * 3 + d[2];
}

// A fragments in lane order, per box: [kFrag1 + kDepth * kFrag2][64
// lanes][8] BF16 (layer 1, then layer 2 head by head), so a lane loads each
// fragment with one 16-byte read (layouts above).
constexpr int kFrag1 = 9, kFrag2 = 18;                    // per box; kFrag2 per layer-2 head
constexpr int kPackBox = (kFrag1 + kDepth * kFrag2) * 64 * 8;   // BF16 values per box
constexpr int kPackFloats = 6 * kPackBox / 2;             // FP32 words of scratch for 6 boxes

__global__ __launch_bounds__(256) void shell_pack(const uint16_t* w, int nbox, ShellTiles tl, uint16_t* packed) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= nbox * kPackBox) return;
    const int b = idx / kPackBox, r = idx % kPackBox;
    const int f = r / 512, lane = (r / 8) % 64, i = r % 8;
    const int g = lane >> 4, n = lane & 15;
    const int* p = tl.perm[b];
    uint16_t v = 0;
    if (f < kFrag1) {
        if (g < 3 && i < kIn) v = w[(size_t)b * kW1 + (n * kIn + i) * kTaps + tap_of(p, f / 3, f % 3, g)];
    } else {
        const int dep = (f - kFrag1) / kFrag2, f2 = (f - kFrag1) % kFrag2;
        const int pair = f2 >> 1, h = f2 & 1;
        const int d2 = h == 0 ? (g >> 1) : 2;
        if (n < kOut && (h == 0 || g < 2))
            v = w[(size_t)nbox * kW1 + ((size_t)b * kDepth + dep) * kW2 + (n * kHid + 8 * (g & 1) + i) * kTaps +
                  tap_of(p, pair / 3, pair % 3, d2)];
    }
    packed[idx] = v;
}

// floor(c / d) for 0 <= c < 4096 and 1 <= d <= 64 from r = 1/d: (c + 0.5)/d
// lies at least 1/(2d) >= 2^-7 from an integer and the product's error is
// below 2^-9.
__device__ __forceinline__ int qdiv(int c, float r) { return (int)(((float)c + 0.5f) * r); }

// periodic wrap of -2 <= c < 3s (a tile's input reaches 17 cells past its
// origin, at most s + 15 < 3s with s >= 8)
__device__ __forceinline__ int wrap2(int c, int s) {
    if (c < 0) return c + s;
    if (c >= s) c -= s;
    return c >= s ? c - s : c;
}

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const uint4& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

constexpr int kInIter = (kInMax + 255) / 256;           // input cells per thread (upper bound)

template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FDTD_SHELL_WAVES))) void shell_mfma(ShellArgs<T> a, ShellTiles tl, const uint16_t* packed, float* ybuf) {
    // the input tile [cell][8 x BF16], then (after layer 1) the hidden tile
    // [2 planes][cell][8 x BF16] in the same words
    __shared__ uint4 s_buf[kTileLds];
    // hardware workgroup k runs on XCD k % 8; each XCD gets a contiguous
    // range of tiles so the halos of neighbouring tiles meet in its L2
    const int nb = gridDim.x, qq = nb >> 3, rr = nb & 7, xcd = blockIdx.x & 7, jj = blockIdx.x >> 3;
    const int tile = (xcd < rr ? xcd * (qq + 1) : rr * (qq + 1) + (xcd - rr) * qq) + jj;
    int b = 0;
    while (b < a.nbox - 1 && tile >= tl.tstart[b + 1]) ++b;
    const ShellBox& B = a.box[b];
    const int p0 = tl.perm[b][0], p1 = tl.perm[b][1], p2 = tl.perm[b][2];
    const int te0 = tl.te[b][0];
    const int ie2 = kTileA2 + 4, he1 = kTileA1 + 2, he0 = te0 + 2;
    const int tix = tile - tl.tstart[b];
    const int t2 = tix % tl.nt[b][2], t1 = (tix / tl.nt[b][2]) % tl.nt[b][1], t0 = tix / (tl.nt[b][2] * tl.nt[b][1]);
    int o[3];                                  // tile origin, grid axes
    o[p0] = B.o0[p0] + t0 * te0;
    o[p1] = B.o0[p1] + t1 * kTileA1;
    o[p2] = B.o0[p2] + t2 * kTileA2;
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    // 1. input tile -> LDS (BF16, 8 per cell), loaded in grid order (z
    // fastest) in passes whose loads are all issued before the first
    // conversion (threads past the tile reload its last cell)
    {
        int ext[3], str[3];                    // input extents and LDS strides per grid axis
        ext[p0] = te0 + 4; ext[p1] = kTileA1 + 4; ext[p2] = ie2;
        str[p0] = kInA0; str[p1] = ie2; str[p2] = 1;
        const int nin = ext[0] * ext[1] * ext[2];
        const float r2 = 1.0f / ext[2], r1 = 1.0f / ext[1];
        const T* E0 = a.E; const T* E1 = a.E + N; const T* E2 = a.E + 2 * N;
        const T* H0 = a.H; const T* H1 = a.H + N; const T* H2 = a.H + 2 * N;
        constexpr int U = sizeof(T) == 4 ? (kInIter + 1) / 2 : 2;   // cells per pass
        for (int c0 = 0; c0 < nin; c0 += U * 256) {
            float v[U][6];
            int dst[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int c = min(c0 + u * 256 + (int)threadIdx.x, nin - 1);
                const int q = qdiv(c, r2), lz = c - q * ext[2];
                const int lx = qdiv(q, r1), ly = q - lx * ext[1];
                dst[u] = lx * str[0] + ly * str[1] + lz * str[2];
                const uint32_t gi = (uint32_t)((wrap2(o[0] - 2 + lx, a.sx) * a.sy + wrap2(o[1] - 2 + ly, a.sy)) * a.sz +
                                               wrap2(o[2] - 2 + lz, a.sz));
                v[u][0] = (float)E0[gi]; v[u][1] = (float)E1[gi]; v[u][2] = (float)E2[gi];
                v[u][3] = (float)H0[gi]; v[u][4] = (float)H1[gi]; v[u][5] = (float)H2[gi];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (c0 + u * 256 + (int)threadIdx.x < nin)
                    s_buf[dst[u]] = make_uint4(pack_bf16(v[u][0], v[u][1]), pack_bf16(v[u][2], v[u][3]),
                                               pack_bf16(v[u][4], v[u][5]), 0u);
            }
        }
    }
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4, n = lane & 15;
    const bf16x8* frag = reinterpret_cast<const bf16x8*>(packed) + (size_t)b * (kFrag1 + kDepth * kFrag2) * 64 + lane;
    __syncthreads();
    // 2. layer 1: wave w takes a0 rows h0 = w/2, w/2 + 2, w/2 + 4 (< T+2) and
    // along a1 the kR1 = (kTileA1 + 2) / 2 rows h1 = kR1 (w % 2) + [0, kR1);
    // B fragment (h0 + d0, y) feeds rows y - d1.  Results stay in registers
    // (BF16, channels 4g .. 4g+3 of cell n) until every wave is done with
    // the input tile.
    constexpr int kR1 = (kTileA1 + 2) / 2, kH0 = 3;
    static_assert(2 * kR1 == kTileA1 + 2, "two waves split the hidden rows along a1");
    uint32_t hreg[kH0][kR1][2];
    {
        bf16x8 W1[kFrag1];
#pragma unroll
        for (int f = 0; f < kFrag1; ++f) W1[f] = frag[f * 64];
        const int h1s = kR1 * (wave & 1);
        const int lofs = n + min(g, 2);        // this lane's cell along a2 (group 3: any valid cell)
#pragma unroll
        for (int k = 0; k < kH0; ++k) {
            const int h0 = (wave >> 1) + 2 * k;
            f32x4 acc[kR1];
#pragma unroll
            for (int r = 0; r < kR1; ++r) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
            if (h0 < he0) {
#pragma unroll
                for (int d0 = 0; d0 < 3; ++d0) {
                    const uint4* rowp = s_buf + (h0 + d0) * kInA0 + h1s * ie2 + lofs;
#pragma unroll
                    for (int y = 0; y < kR1 + 2; ++y) {
                        const uint4 bv = rowp[y * ie2];
#pragma unroll
                        for (int d1 = 0; d1 < 3; ++d1) {
                            const int r = y - d1;
                            if (r >= 0 && r < kR1) acc[r] = mfma(W1[d0 * 3 + d1], bv, acc[r]);
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < kR1; ++r) {
                hreg[k][r][0] = pack_bf16(acc[r][0], acc[r][1]);
                hreg[k][r][1] = pack_bf16(acc[r][2], acc[r][3]);
            }
        }
    }
    __syncthreads();
    {
        // hidden cell (h0, h1, n) channels 4g .. 4g+3 -> plane g / 2, half g % 2
        char* hb = reinterpret_cast<char*>(s_buf) + (size_t)(g >> 1) * kHidPlane * 16 + n * 16 + (g & 1) * 8;
        const int h1s = kR1 * (wave & 1);
#pragma unroll
        for (int k = 0; k < kH0; ++k) {
            const int h0 = (wave >> 1) + 2 * k;
            if (h0 >= he0) continue;
#pragma unroll
            for (int r = 0; r < kR1; ++r)
                *reinterpret_cast<uint2*>(hb + ((h0 * he1 + h1s + r) * kRow) * 16) = make_uint2(hreg[k][r][0], hreg[k][r][1]);
        }
    }
    __syncthreads();
    // 3. layer 2: wave w takes output a0 row o0 = w (< T), i.e. depth o0 of
    // the box (a tile is the box's whole depth) and its head, and its 8 rows
    // along a1; B fragment (o0 + d0, y, half) feeds rows y - d1.  Lanes 14,
    // 15 of each MFMA read a valid cell and their columns are never used.
    const int o0 = wave;
    if (o0 < te0) {
        bf16x8 W2[kFrag2];                     // rows n >= 3 are zero: only 3 lanes of 16 load
#pragma unroll
        for (int f = 0; f < kFrag2; ++f) {
            const uint4 v = n < kOut ? reinterpret_cast<const uint4*>(frag)[(kFrag1 + o0 * kFrag2 + f) * 64]
                                     : make_uint4(0u, 0u, 0u, 0u);
            W2[f] = __builtin_bit_cast(bf16x8, v);
        }
        constexpr int kR2 = kTileA1;
        f32x4 acc[kR2];
#pragma unroll
        for (int r = 0; r < kR2; ++r) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int nn = min(n, kTileA2 - 1);
        const uint4* plane = s_buf + (g & 1) * kHidPlane + nn;
#pragma unroll
        for (int d0 = 0; d0 < 3; ++d0) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int d2 = h == 0 ? (g >> 1) : 2;
                const uint4* rowp = plane + (o0 + d0) * he1 * kRow + d2;
#pragma unroll
                for (int y = 0; y < kR2 + 2; ++y) {
                    const uint4 bv = rowp[y * kRow];
#pragma unroll
                    for (int d1 = 0; d1 < 3; ++d1) {
                        const int r = y - d1;
                        if (r >= 0 && r < kR2) acc[r] = mfma(W2[(d0 * 3 + d1) * 2 + h], bv, acc[r]);
                    }
                }
            }
        }
        // outputs 0..2 of cell (o0, r, n) (tile axes) to the correction, in
        // box order: ycorr[component][ostart[b] + box-local C index]
        int ext[3], bst[3];
        for (int q = 0; q < 3; ++q) ext[q] = B.o1[q] - B.o0[q];
        bst[0] = ext[1] * ext[2]; bst[1] = ext[2]; bst[2] = 1;
        const int l0 = o[p0] - B.o0[p0] + o0, l1 = o[p1] - B.o0[p1], l2 = o[p2] - B.o0[p2] + n;
        if (g == 0 && n < kTileA2 && l0 < ext[p0] && l2 < ext[p2]) {
            const int out = tl.ostart[a.nbox];
            float* yc = ybuf + tl.ostart[b] + l0 * bst[p0] + l1 * bst[p1] + l2 * bst[p2];
#pragma unroll
            for (int r = 0; r < kR2; ++r) {
                if (l1 + r < ext[p1]) {
                    yc[r * bst[p1]] = acc[r][0];
                    yc[out + r * bst[p1]] = acc[r][1];
                    yc[2 * out + r * bst[p1]] = acc[r][2];
                }
            }
        }
    }
}

// E += correction on the shell cells, one thread per cell in box order
// (z fastest within a box), the order of the correction's scratch.
template <class T>
__global__ __launch_bounds__(256) void shell_add(ShellArgs<T> a, ShellTiles tl, const float* ycorr) {
    int cell = blockIdx.x * 256 + threadIdx.x;
    const int out = tl.ostart[a.nbox];
    if (cell >= out) return;
    const float y0 = ycorr[cell], y1 = ycorr[out + cell], y2 = ycorr[2 * out + cell];
    int b = 0;
    while (b < a.nbox - 1 && cell >= tl.ostart[b + 1]) ++b;
    cell -= tl.ostart[b];
    const ShellBox& B = a.box[b];
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    const int lz = cell % nz, ly = (cell / nz) % ny, lx = cell / (nz * ny);
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const size_t gi = ((size_t)(B.o0[0] + lx) * a.sy + (B.o0[1] + ly)) * a.sz + (B.o0[2] + lz);
    a.Eout[gi] += (T)y0;
    a.Eout[N + gi] += (T)y1;
    a.Eout[2 * N + gi] += (T)y2;
}

struct ShellPlan {
    int nbox;
    ShellBox box[6];
    int hcells, out_cells;
    ShellTiles tiles;
    int ntiles;
    size_t scratch_floats;   // hidden values (backward) or tile-major correction (forward)
};

ShellPlan make_shell(const Plan& P, int T) {
    ShellPlan S;
    const int s[3] = {P.sx, P.sy, P.sz};
    int nb = 0;
    for (int r = 0; r < 3; ++r) {
        for (int side = 0; side < 2; ++side) {
            ShellBox& B = S.box[nb++];
            for (int q = 0; q < 3; ++q) {
                if (q < r) { B.o0[q] = T; B.o1[q] = s[q] - T; }
                else if (q > r) { B.o0[q] = 0; B.o1[q] = s[q]; }
                else { B.o0[q] = side == 0 ? 0 : s[q] - T; B.o1[q] = side == 0 ? T : s[q]; }
            }
        }
    }
    S.nbox = nb;
    int off = 0, outc = 0;
    for (int b = 0; b < nb; ++b) {
        ShellBox& B = S.box[b];
        for (int q = 0; q < 3; ++q) {
            B.s[q] = s[q];
            B.full[q] = (B.o0[q] == 0 && B.o1[q] == s[q]) ? 1 : 0;
            B.e[q] = B.full[q] ? s[q] : B.o1[q] - B.o0[q] + 2;
        }
        B.hcells = B.e[0] * B.e[1] * B.e[2];
        B.hoff = off;
        off += B.hcells;
        outc += (B.o1[0] - B.o0[0]) * (B.o1[1] - B.o0[1]) * (B.o1[2] - B.o0[2]);
    }
    S.hcells = off;
    S.out_cells = outc;
    // MFMA tiles: a0 = the box's thin axis (T), a2 = z for x and y faces, y
    // for z faces (14), a1 = the remaining axis (8)
    int ts = 0;
    for (int b = 0; b < nb; ++b) {
        const int r = b / 2;
        const int a2 = (r == 2) ? 1 : 2;
        const int a1 = 3 - r - a2;
        const int perm[3] = {r, a1, a2};
        const int te[3] = {T, kTileA1, kTileA2};
        for (int q = 0; q < 3; ++q) {
            const int ext = S.box[b].o1[perm[q]] - S.box[b].o0[perm[q]];   // T along a0: one tile
            S.tiles.perm[b][q] = perm[q];
            S.tiles.te[b][q] = te[q];
            S.tiles.nt[b][q] = (ext + te[q] - 1) / te[q];
        }
        S.tiles.tstart[b] = ts;
        ts += S.tiles.nt[b][0] * S.tiles.nt[b][1] * S.tiles.nt[b][2];
    }
    int oc = 0;
    for (int b = 0; b < nb; ++b) {
        S.tiles.ostart[b] = oc;
        oc += (S.box[b].o1[0] - S.box[b].o0[0]) * (S.box[b].o1[1] - S.box[b].o0[1]) * (S.box[b].o1[2] - S.box[b].o0[2]);
    }
    S.tiles.ostart[nb] = oc;
    S.tiles.tstart[nb] = ts;
    S.ntiles = ts;
    S.scratch_floats = std::max((size_t)S.hcells * kHid, (size_t)3 * S.out_cells);
    return S;
}

template <class T>
void fill_shell(const Plan& P, const ShellPlan& S, ShellArgs<T>& a) {
    memset(&a, 0, sizeof(a));
    a.sx = P.sx; a.sy = P.sy; a.sz = P.sz;
    a.nbox = S.nbox;
    for (int b = 0; b < S.nbox; ++b) a.box[b] = S.box[b];
}

// T >= 2 keeps every grown region free of self-overlap through the wrap;
// T <= 4 bounds the MFMA tiles' LDS.
int check_shell(const Plan& P, int T) {
    if (T < 2 || T > kMaxT) return fail(3, "shell thickness %d not in [2, %d]", T, kMaxT);
    if (2 * T + 4 > P.sx || 2 * T + 4 > P.sy || 2 * T + 4 > P.sz)
        return fail(3, "shell thickness %d too large for the grid", T);
    return 0;
}

// Scratch layout: [0, scratch_floats) the correction, [3][out_cells] FP32
// in box order, between its computation and its addition to E (the backward
// keeps hidden values there), then the packed weight fragments.
uint16_t* packed_weights(const ShellPlan& S, float* scratch) {
    return reinterpret_cast<uint16_t*>(scratch + S.scratch_floats);
}

void launch_shell_pack(const ShellPlan& S, const uint16_t* w, float* scratch, hipStream_t st) {
    hipLaunchKernelGGL(shell_pack, dim3((S.nbox * kPackBox + 255) / 256), dim3(256), 0, st, w, S.nbox, S.tiles,
                       packed_weights(S, scratch));
}

// The correction (weights already packed in scratch) into the scratch, and
// its addition to E: two launches, so that a caller can run something else
// between them.
template <class T>
ShellArgs<T> shell_args(const Plan& P, const ShellPlan& S, const uint16_t* w, const void* E, const void* H) {
    ShellArgs<T> a;
    fill_shell<T>(P, S, a);
    a.w = w; a.E = (const T*)E; a.H = (const T*)H; a.Eout = (T*)E;
    return a;
}

void launch_shell_mfma(const Plan& P, const ShellPlan& S, const uint16_t* w, const void* E, const void* H,
                       float* scratch, hipStream_t st) {
    const uint16_t* pk = packed_weights(S, scratch);
    if (P.d.dtype == FDTD_F32)
        hipLaunchKernelGGL(shell_mfma<float>, dim3(S.ntiles), dim3(256), 0, st, shell_args<float>(P, S, w, E, H),
                           S.tiles, pk, scratch);
    else
        hipLaunchKernelGGL(shell_mfma<double>, dim3(S.ntiles), dim3(256), 0, st, shell_args<double>(P, S, w, E, H),
                           S.tiles, pk, scratch);
}

void launch_shell_add(const Plan& P, const ShellPlan& S, void* E, const float* scratch, hipStream_t st) {
    const dim3 grid((S.out_cells + 255) / 256);
    if (P.d.dtype == FDTD_F32)
        hipLaunchKernelGGL(shell_add<float>, grid, dim3(256), 0, st, shell_args<float>(P, S, nullptr, E, nullptr),
                           S.tiles, scratch);
    else
        hipLaunchKernelGGL(shell_add<double>, grid, dim3(256), 0, st, shell_args<double>(P, S, nullptr, E, nullptr),
                           S.tiles, scratch);
}

}  // namespace

extern "C" {

int64_t fdtd_shell_hidden_elems(const void* plan, int32_t thickness) {
    if (!plan) return -1;
    const Plan& P = *(const Plan*)plan;
    if (check_shell(P, thickness)) return -1;
    return (int64_t)make_shell(P, thickness).scratch_floats + kPackFloats;
}

int64_t fdtd_shell_weight_elems(void) { return 6 * (int64_t)(kW1 + kDepth * kW2); }

int fdtd_shell_apply(const void* plan, int32_t thickness, const uint16_t* w, void* E, const void* H,
                     float* hidden, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_shell(P, thickness)) return rc;
    if (!w || !E || !H || !hidden) return fail(2, "w, E, H and hidden are required");
    const ShellPlan S = make_shell(P, thickness);
    hipStream_t st = (hipStream_t)stream;
    launch_shell_pack(S, w, hidden, st);
    launch_shell_mfma(P, S, w, E, H, hidden, st);
    launch_shell_add(P, S, E, hidden, st);
    return check_launch("fdtd_shell_apply");
}

int fdtd_shell_step_n(const void* plan, const fdtd_fields* f, int32_t thickness, const uint16_t* w, float* hidden,
                      int32_t n, int64_t src_index, const void* src_wave, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_shell(P, thickness)) return rc;
    if (!w || !hidden) return fail(2, "w and hidden are required");
    if (n < 0) return fail(3, "n = %d < 0", n);
    if (src_wave && (src_index < 0 || src_index >= 3 * P.N))
        return fail(3, "src_index %lld out of [0, %lld)", (long long)src_index, (long long)(3 * P.N));
    if (P.d.poynting == FDTD_POYNTING_TWO_PASS) return fail(3, "two-pass Poynting is not supported with the shell");
    if (P.d.bloch_mask) return fail(3, "the shell correction is for real fields (no Bloch boundaries)");
    if (int rc = check_fields(P, f, true)) return rc;
    const ShellPlan S = make_shell(P, thickness);
    hipStream_t st = (hipStream_t)stream;
    fdtd_fields cur = *f;
    const size_t esz = P.d.dtype == FDTD_F32 ? 4 : 8;
    const bool swap = (f->H_out && f->H_out != f->H);
    if (n > 0) launch_shell_pack(S, w, hidden, st);
    // step k: E update (+ the correction of step k-1, added before the
    // update) and source; correction y(E, H) into the scratch; H update from
    // the uncorrected E.  The last step's correction is added at the end.
    const auto amp_of = [&](int k) -> const void* {
        return src_wave ? (const char*)src_wave + (size_t)k * esz : nullptr;
    };
    const auto half_e = [&](int k) {
        const float* pending = k > 0 ? hidden : nullptr;
        return (P.d.dtype == FDTD_F32)
                   ? launch_e<float>(P, &cur, st, (const float*)amp_of(k), src_index, pending, thickness, S.out_cells)
                   : launch_e<double>(P, &cur, st, (const double*)amp_of(k), src_index, pending, thickness,
                                      S.out_cells);
    };
    const auto half_h = [&]() {
        const int rc = (P.d.dtype == FDTD_F32) ? launch_h<float>(P, &cur, st) : launch_h<double>(P, &cur, st);
        if (swap) {
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
        }
        return rc;
    };
    // With a second E buffer and a plan that has the fused pass: E_0 and its
    // correction, then per step one pass for H_{k-1} + S and E_k (+ the
    // correction of step k-1) and the correction of step k, then H_{n-1}.
    // Same order of operations per cell as below; the count of passes is
    // kept even so that E ends in f->E.
    if (n >= 2 && f->E_out && f->E_out != f->E && swap && he_supported(P)) {
        if (int rc = half_e(0)) return rc;
        launch_shell_mfma(P, S, w, cur.E, cur.H, hidden, st);
        int k = 1;
        if ((n - 1) % 2 == 1) {
            if (int rc = half_h()) return rc;
            if (int rc = half_e(1)) return rc;
            launch_shell_mfma(P, S, w, cur.E, cur.H, hidden, st);
            k = 2;
        }
        for (; k < n; ++k) {
            const int rc = (P.d.dtype == FDTD_F32)
                               ? launch_he<float>(P, &cur, st, (const float*)amp_of(k), src_index, hidden,
                                                  thickness, S.out_cells)
                               : launch_he<double>(P, &cur, st, (const double*)amp_of(k), src_index, hidden,
                                                   thickness, S.out_cells);
            if (rc) return rc;
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
            std::swap(cur.E, cur.E_out);
            launch_shell_mfma(P, S, w, cur.E, cur.H, hidden, st);
        }
        if (int rc = half_h()) return rc;
        launch_shell_add(P, S, cur.E, hidden, st);
        return check_launch("fdtd_shell_step_n");
    }
    for (int k = 0; k < n; ++k) {
        const void* amp = src_wave ? (const char*)src_wave + (size_t)k * esz : nullptr;
        const float* pending = k > 0 ? hidden : nullptr;
        int rc = (P.d.dtype == FDTD_F32)
                     ? launch_e<float>(P, &cur, st, (const float*)amp, src_index, pending, thickness, S.out_cells)
                     : launch_e<double>(P, &cur, st, (const double*)amp, src_index, pending, thickness, S.out_cells);
        if (rc) return rc;
        launch_shell_mfma(P, S, w, cur.E, cur.H, hidden, st);
        rc = (P.d.dtype == FDTD_F32) ? launch_h<float>(P, &cur, st) : launch_h<double>(P, &cur, st);
        if (rc) return rc;
        if (swap) {
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
        }
    }
    if (n > 0) launch_shell_add(P, S, cur.E, hidden, st);
    return check_launch("fdtd_shell_step_n");
}

int64_t fdtd_shell_partials_elems(void) { return (int64_t)64 * std::max(kW1, kDepth * kW2); }

int fdtd_shell_backward(const void* plan, int32_t thickness, const uint16_t* w, const void* E, const void* H,
                        float* hidden, float* hbar, const void* Ebar, void* Ebar_out, void* Hbar_out,
                        float* partials, float* wgrad, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_shell(P, thickness)) return rc;
    if (!w || !E || !H || !hidden || !hbar || !Ebar || !Ebar_out || !Hbar_out || !partials || !wgrad)
        return fail(2, "all buffers are required");
    if (Ebar == Ebar_out) return fail(2, "Ebar_out must differ from Ebar");
    const ShellPlan S = make_shell(P, thickness);
    hipStream_t st = (hipStream_t)stream;
    const int groups = 64;
    auto run = [&](auto tag) -> int {
        using T = decltype(tag);
        ShellArgs<T> a;
        fill_shell<T>(P, S, a);
        a.w = w; a.E = (const T*)E; a.H = (const T*)H; a.hid = hidden; a.hbar = hbar;
        a.Ebar = (const T*)Ebar; a.Ebar_out = (T*)Ebar_out; a.Hbar_out = (T*)Hbar_out;
        a.wgrad_part = partials;
        const int g1 = (S.hcells + 255) / 256;
        // forward hidden (the weight gradient of layer 2 needs it)
        hipLaunchKernelGGL(shell_l1<T>, dim3(g1), dim3(256), 0, st, a, S.hcells);
        hipLaunchKernelGGL(shell_l2_t<T>, dim3(g1), dim3(256), 0, st, a, S.hcells);
        for (int b = 0; b < S.nbox; ++b) {
            a.part_box = b;
            const ShellBox& B = S.box[b];
            const int n2 = (B.full[0] ? B.s[0] : B.o1[0] - B.o0[0] + 4) *
                           (B.full[1] ? B.s[1] : B.o1[1] - B.o0[1] + 4) *
                           (B.full[2] ? B.s[2] : B.o1[2] - B.o0[2] + 4);
            hipLaunchKernelGGL(shell_l1_t<T>, dim3((n2 + 255) / 256), dim3(256), 0, st, a);
            // W1 gradient of box b
            hipLaunchKernelGGL(shell_wgrad1<T>, dim3(groups), dim3(256), 0, st, a);
            hipLaunchKernelGGL(shell_wgrad_sum, dim3((kW1 + 255) / 256), dim3(256), 0, st, (const float*)partials,
                               groups, kW1, wgrad + (size_t)b * kW1);
            // W2 gradient of box b
            hipLaunchKernelGGL(shell_wgrad2<T>, dim3(groups), dim3(256), 0, st, a);
            hipLaunchKernelGGL(shell_wgrad_sum, dim3((kDepth * kW2 + 255) / 256), dim3(256), 0, st,
                               (const float*)partials, groups, kDepth * kW2,
                               wgrad + (size_t)S.nbox * kW1 + (size_t)b * kDepth * kW2);
        }
        return 0;
    };
    if (P.d.dtype == FDTD_F32) run(0.f);
    else run(0.0);
    return check_launch("fdtd_shell_backward");
}

}  // extern "C"
