// The learned absorbing shell (part of the fdtd_hip.hip translation unit).
//
// A thin shell of T cells on every face runs the CPML-T update; after each E
// update (and source) a small convolutional net adds a correction to E on the
// shell cells:
//   x = (Ex, Ey, Ez, Hx, Hy, Hz)                         6 input channels
//   h_k(j) = sum_{c,t} W1[k][c][t] x_c(j + t)            16 hidden channels
//   y_a(i) = sum_{k,t} W2[a][k][t] h_k(i + t)            3 outputs
//   E_a(i) += y_a(i)  for i in the shell
// t runs over the 27 offsets of a 3x3x3 stencil (periodic wrap, as the step).
// There is no activation: the correction must be linear in the fields, as
// Maxwell's equations are.  Weights and activations are BF16 (inputs rounded
// to BF16), products are exact in FP32 and accumulated in FP32.
//
// The shell is split into 6 boxes, one per face (x faces own the edges and
// corners they touch, then y faces), each with its own weights.  Hidden values
// live per box on the box grown by one cell, in a caller-provided FP32
// scratch [cell][16].

namespace {

constexpr int kHid = 16;
constexpr int kIn = 6;
constexpr int kOut = 3;
constexpr int kTaps = 27;
constexpr int kW1 = kHid * kIn * kTaps;    // per box
constexpr int kW2 = kOut * kHid * kTaps;   // per box

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

// layer 2 over the output cells, adding the correction to Eout
template <class T>
__global__ __launch_bounds__(256) void shell_l2(ShellArgs<T> a, int total_out) {
    int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= total_out) return;
    int b = 0;
    for (; b < a.nbox; ++b) {
        const ShellBox& B = a.box[b];
        const int n = (B.o1[0] - B.o0[0]) * (B.o1[1] - B.o0[1]) * (B.o1[2] - B.o0[2]);
        if (cell < n) break;
        cell -= n;
    }
    if (b >= a.nbox) return;
    const ShellBox& B = a.box[b];
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    const int lz = cell % nz, ly = (cell / nz) % ny, lx = cell / (nz * ny);
    const uint16_t* w2 = a.w + (size_t)a.nbox * kW1 + (size_t)b * kW2;
    float y[kOut] = {0.f, 0.f, 0.f};
    for (int t = 0; t < kTaps; ++t) {
        const int hx = hid_local(B, 0, B.o0[0] + lx + t / 9 - 1);
        const int hy = hid_local(B, 1, B.o0[1] + ly + (t / 3) % 3 - 1);
        const int hz = hid_local(B, 2, B.o0[2] + lz + t % 3 - 1);
        const float* h = a.hid + ((size_t)B.hoff + ((size_t)hx * B.e[1] + hy) * B.e[2] + hz) * kHid;
#pragma unroll
        for (int k = 0; k < kHid; ++k) {
            const float hv = h[k];
#pragma unroll
            for (int o = 0; o < kOut; ++o) y[o] += bf16_to_f(w2[(o * kHid + k) * kTaps + t]) * hv;
        }
    }
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const size_t gi = gidx(a, B.o0[0] + lx, B.o0[1] + ly, B.o0[2] + lz);
#pragma unroll
    for (int o = 0; o < kOut; ++o) a.Eout[o * N + gi] += (T)y[o];
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
    const uint16_t* w2 = a.w + (size_t)a.nbox * kW1 + (size_t)b * kW2;
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

template <class T>
__global__ __launch_bounds__(256) void shell_wgrad2(ShellArgs<T> a) {
    __shared__ float sy[kChunk2][kOut];
    __shared__ float shc[kChunk2][kHid * kTaps];
    const ShellBox& B = a.box[a.part_box];
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    const int nout = (B.o1[0] - B.o0[0]) * ny * nz;
    const int per = (nout + gridDim.x - 1) / gridDim.x;
    const int c0 = blockIdx.x * per, c1 = min(c0 + per, nout);
    constexpr int NW = kOut * kHid * kTaps;  // 1296
    constexpr int PER = (NW + 255) / 256;
    float acc[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) acc[m] = 0.f;
    for (int base = c0; base < c1; base += kChunk2) {
        const int n = min(kChunk2, c1 - base);
        for (int e = threadIdx.x; e < kChunk2 * kOut; e += 256) {
            const int l = e / kOut, o = e % kOut;
            float v = 0.f;
            if (l < n) {
                const int cl = base + l;
                const int lz = cl % nz, ly = (cl / nz) % ny, lx = cl / (nz * ny);
                v = (float)a.Ebar[o * N + gidx(a, B.o0[0] + lx, B.o0[1] + ly, B.o0[2] + lz)];
            }
            sy[l][o] = v;
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
                const int o = wi / (kHid * kTaps), kt = wi % (kHid * kTaps);
                float s_ = 0.f;
                for (int l = 0; l < kChunk2; ++l) s_ += sy[l][o] * shc[l][kt];
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
// layer grown by 1 (T+2 x 10 x 16) as BF16 in two planes [8 channels][cell]
// over the same LDS words.  Both layers are GEMMs over (tap, channel), with
// M = 16 cells of one a2 row (a hidden row is exactly 16 cells; an output
// row is 14, the last 2 rows of each MFMA are discarded):
//   layer 1: K = 28 taps x 8 channels (tap 27 and channels 6, 7 are zero),
//            N = 16 hidden channels;
//   layer 2: K = 28 taps x 16 hidden channels, N = 16 (3 outputs used).
// MFMA operands (16x16x32): lane l holds A row l%16 and B column l%16 at
// k = 8*(l/16) + [0..7]; D column l%16 at rows 4*(l/16) + [0..3].  With the
// K order above, a lane's 8 A values are one 16-byte LDS read at a per-lane
// offset (fixed per fragment) from the row's base.  The correction goes to
// a tile-major scratch [tile][3][T*8*16] and is added to E by shell_add,
// since neighbouring tiles read E around their own cells.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxT = 4;
constexpr int kTileA1 = 8, kTileA2 = 14;                 // output cells along a1, a2
constexpr int kRow = 16;                                  // MFMA rows = hidden a2 extent
constexpr int kInMax = (kMaxT + 4) * (kTileA1 + 4) * (kTileA2 + 4);
constexpr int kHidRows = (kMaxT + 2) * (kTileA1 + 2);     // hidden a2 rows (of 16 cells)
constexpr int kHidPlane = kHidRows * kRow;                // cells per hidden plane
constexpr int kTileRows = kMaxT * kTileA1 * kRow;         // correction slots per tile and output
constexpr int kTileLds = kInMax > 2 * kHidPlane ? kInMax : 2 * kHidPlane;   // uint4 words
static_assert(kTileA2 + 2 == kRow, "hidden rows must be one MFMA tall");

struct ShellTiles {
    int perm[6][3];    // grid axis of tile axes a0, a1, a2
    int te[6][3];      // output extents along a0, a1, a2
    int nt[6][3];      // tiles along a0, a1, a2
    int tstart[7];     // prefix of tiles per box
};

// BF16 conversions with round-to-nearest-even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    bf16x2 r;
    r[0] = (__bf16)lo;
    r[1] = (__bf16)hi;
    return __builtin_bit_cast(uint32_t, r);
}

// B fragments of both layers in lane order, per box: [kFrag1 + kFrag2][64
// lanes][8] BF16, so a lane loads each fragment with one 16-byte read.
// Layer 1 fragment j holds, for lane (g, n), W1[n][i][4j + g] for input
// channel i < 6 (zero for i = 6, 7 and tap 27); layer 2 fragment j holds
// W2[n][8(g & 1) + i][2j + g / 2] for output n < 3 (zero otherwise).
constexpr int kFrag1 = 7, kFrag2 = 14;
constexpr int kPackBox = (kFrag1 + kFrag2) * 64 * 8;      // BF16 values per box
constexpr int kPackFloats = 6 * kPackBox / 2;             // FP32 words of scratch for 6 boxes

__global__ __launch_bounds__(256) void shell_pack(const uint16_t* w, int nbox, uint16_t* packed) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= nbox * kPackBox) return;
    const int b = idx / kPackBox, r = idx % kPackBox;
    const int j = r / 512, lane = (r / 8) % 64, i = r % 8;
    const int g = lane >> 4, n = lane & 15;
    uint16_t v = 0;
    if (j < kFrag1) {
        const int tap = 4 * j + g;
        if (tap < kTaps && i < kIn) v = w[(size_t)b * kW1 + (n * kIn + i) * kTaps + tap];
    } else {
        const int tap = 2 * (j - kFrag1) + (g >> 1);
        const int k = 8 * (g & 1) + i;
        if (tap < kTaps && n < kOut) v = w[(size_t)nbox * kW1 + (size_t)b * kW2 + (n * kHid + k) * kTaps + tap];
    }
    packed[idx] = v;
}

// floor(c / d) for 0 <= c < 4096 and 1 <= d <= 64 from r = 1/d: (c + 0.5)/d
// lies at least 1/(2d) >= 2^-7 from an integer and the product's error is
// below 2^-9.
__device__ __forceinline__ int qdiv(int c, float r) { return (int)(((float)c + 0.5f) * r); }

// displacement along tile axis a of stencil tap t (offsets 0..2 per grid axis)
__device__ __forceinline__ int tap_disp(int t, int grid_axis) {
    return grid_axis == 0 ? t / 9 : (grid_axis == 1 ? (t / 3) % 3 : t % 3);
}

// periodic wrap of -2 <= c < 3s (a tile's input reaches 17 cells past its
// origin, at most s + 15 < 3s with s >= 8)
__device__ __forceinline__ int wrap2(int c, int s) {
    if (c < 0) return c + s;
    if (c >= s) c -= s;
    return c >= s ? c - s : c;
}

constexpr int kInIter = (kInMax + 255) / 256;           // input cells per thread
constexpr int kHidPerWave = (kHidRows + 3) / 4;         // hidden rows per wave

template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void shell_mfma(ShellArgs<T> a, ShellTiles tl, const uint16_t* packed, float* ybuf,
                                                  int per_block) {
    // the input tile [cell][8 x BF16], then (after layer 1) the hidden tile
    // [2 planes][cell][8 x BF16] in the same words
    __shared__ uint4 s_buf[kTileLds];
    // workgroup -> a run of per_block consecutive tiles; hardware workgroup
    // k runs on XCD k % 8, and each XCD gets a contiguous range of runs so
    // the halos of neighbouring tiles meet in its L2
    const int nb = gridDim.x, qq = nb >> 3, rr = nb & 7, xcd = blockIdx.x & 7, jj = blockIdx.x >> 3;
    const int run = (xcd < rr ? xcd * (qq + 1) : rr * (qq + 1) + (xcd - rr) * qq) + jj;
    const int first = run * per_block, last = min(first + per_block, tl.tstart[a.nbox]);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4, n = lane & 15;
    for (int tile = first; tile < last; ++tile) {
    int b = 0;
    while (b < a.nbox - 1 && tile >= tl.tstart[b + 1]) ++b;
    const ShellBox& B = a.box[b];
    const int p0 = tl.perm[b][0], p1 = tl.perm[b][1], p2 = tl.perm[b][2];
    const int te0 = tl.te[b][0], te1 = tl.te[b][1];
    const int ie1 = te1 + 4, ie2 = kTileA2 + 4;
    const int he0 = te0 + 2, he1 = te1 + 2;
    const int tix = tile - tl.tstart[b];
    const int t2 = tix % tl.nt[b][2], t1 = (tix / tl.nt[b][2]) % tl.nt[b][1], t0 = tix / (tl.nt[b][2] * tl.nt[b][1]);
    int o[3];                                  // tile origin, grid axes
    o[p0] = B.o0[p0] + t0 * te0;
    o[p1] = B.o0[p1] + t1 * te1;
    o[p2] = B.o0[p2] + t2 * kTileA2;
    const bf16x8* frag = reinterpret_cast<const bf16x8*>(packed) + (size_t)b * (kFrag1 + kFrag2) * 64 + lane;
    if (tile > first) __syncthreads();        // the previous tile's layer 2 is done with s_buf
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    // 1. input tile -> LDS (BF16, 8 per cell), loaded in grid order (z
    // fastest); every load of a pass is issued before the first conversion
    // (threads past the tile reload its last cell)
    {
        int ext[3], str[3];                    // input extents and LDS strides per grid axis
        ext[p0] = te0 + 4; ext[p1] = ie1; ext[p2] = ie2;
        str[p0] = ie1 * ie2; str[p1] = ie2; str[p2] = 1;
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
    bf16x8 B1[kFrag1];
#pragma unroll
    for (int j = 0; j < kFrag1; ++j) B1[j] = frag[j * 64];
    __syncthreads();
    // 2. layer 1: hidden row (h0, h1) = 16 cells along a2; a wave keeps its
    // rows (BF16) in registers until every wave is done with the input tile
    const int mt1 = he0 * he1;
    uint32_t hreg[kHidPerWave][2];
    {
        int off1[kFrag1];                      // LDS word of this lane's input cell, from the row base
#pragma unroll
        for (int j = 0; j < kFrag1; ++j) {
            const int tap = min(4 * j + g, kTaps - 1);     // tap 27 has zero weights
            off1[j] = (tap_disp(tap, p0) * ie1 + tap_disp(tap, p1)) * ie2 + tap_disp(tap, p2) + n;
        }
        // rows wave + 4i; past the last row (T < 4) a wave repeats the last
        // row, whose values it then writes again unchanged
        int h0 = 0, h1 = wave;                 // row wave + 4i = h0 * he1 + h1 (he1 >= 10 > 4)
#pragma unroll
        for (int i = 0; i < kHidPerWave; ++i) {
            if (i > 0) {
                h1 += 4;
                if (h1 >= he1) { h1 -= he1; ++h0; }
            }
            const int rh0 = h0 < he0 ? h0 : he0 - 1, rh1 = h0 < he0 ? h1 : he1 - 1;
            const uint4* base = s_buf + (rh0 * ie1 + rh1) * ie2;
            uint4 av[kFrag1];                  // every read of the row in flight before the first MFMA
#pragma unroll
            for (int j = 0; j < kFrag1; ++j) av[j] = base[off1[j]];
            __builtin_amdgcn_sched_barrier(0);    // keeps the reads ahead of the MFMA chain
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int j = 0; j < kFrag1; ++j)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[j]), B1[j], acc, 0, 0, 0);
            hreg[i][0] = pack_bf16(acc[0], acc[1]);
            hreg[i][1] = pack_bf16(acc[2], acc[3]);
        }
    }
    __syncthreads();
    {
        // cell (row mt, a2 = 4g + q), channel n -> plane n / 8, element n % 8
        uint16_t* hid16 = reinterpret_cast<uint16_t*>(s_buf) + ((n >> 3) * kHidPlane + 4 * g) * 8 + (n & 7);
#pragma unroll
        for (int i = 0; i < kHidPerWave; ++i) {
            const int mt = min(wave + 4 * i, mt1 - 1);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                hid16[(mt * kRow + q) * 8] = (uint16_t)(hreg[i][q >> 1] >> (16 * (q & 1)));
        }
    }
    __syncthreads();
    // 3. layer 2: output row (o0, o1) = 14 cells along a2 (lanes 14, 15 of
    // each MFMA read a valid cell and their rows are never used)
    bf16x8 B2[kFrag2];                         // columns n >= 3 are zero: only 3 lanes of 16 load
#pragma unroll
    for (int j = 0; j < kFrag2; ++j) {
        const uint4 v = n < kOut ? reinterpret_cast<const uint4*>(frag)[(kFrag1 + j) * 64] : make_uint4(0u, 0u, 0u, 0u);
        B2[j] = __builtin_bit_cast(bf16x8, v);
    }
    int off2[kFrag2];
    {
        const int nn = min(n, kTileA2 - 1);
#pragma unroll
        for (int j = 0; j < kFrag2; ++j) {
            const int tap = min(2 * j + (g >> 1), kTaps - 1);
            off2[j] = (tap_disp(tap, p0) * he1 + tap_disp(tap, p1)) * kRow + tap_disp(tap, p2) + nn +
                      (g & 1) * kHidPlane;
        }
    }
    const int mt2 = te0 * te1;
    float* yt = ybuf + (size_t)tile * kOut * kTileRows + 4 * g;
    int r0 = 0, r1 = wave;                     // row mt = r0 * te1 + r1 (te1 = 8 > 4)
    for (int mt = wave; mt < mt2; mt += 4, r1 += 4) {
        if (r1 >= te1) { r1 -= te1; ++r0; }
        const uint4* base = s_buf + (r0 * he1 + r1) * kRow;
        uint4 av[kFrag2];
#pragma unroll
        for (int j = 0; j < kFrag2; ++j) av[j] = base[off2[j]];
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < kFrag2; ++j)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[j]), B2[j], acc, 0, 0, 0);
        if (n < kOut)      // slots 4g .. 4g+3 of row mt, output n
            *reinterpret_cast<f32x4*>(yt + (size_t)n * kTileRows + mt * kRow) = acc;
    }
    }
}

// E += correction on the shell cells: box cell l (grid axes) sits in tile
// (l[p_a] / te_a) at slot ((l[p0] % te0) * te1 + l[p1] % te1) * 16 + l[p2] % te2
template <class T>
__global__ __launch_bounds__(256) void shell_add(ShellArgs<T> a, ShellTiles tl, const float* ybuf, int total_out) {
    int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= total_out) return;
    int b = 0;
    for (; b < a.nbox; ++b) {
        const ShellBox& B = a.box[b];
        const int n = (B.o1[0] - B.o0[0]) * (B.o1[1] - B.o0[1]) * (B.o1[2] - B.o0[2]);
        if (cell < n) break;
        cell -= n;
    }
    if (b >= a.nbox) return;
    const ShellBox& B = a.box[b];
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    int l[3];
    l[2] = cell % nz;
    l[1] = (cell / nz) % ny;
    l[0] = cell / (nz * ny);
    const int u0 = l[tl.perm[b][0]], u1 = l[tl.perm[b][1]], u2 = l[tl.perm[b][2]];
    const int t0 = u0 / tl.te[b][0], t1 = u1 / tl.te[b][1], t2 = u2 / tl.te[b][2];
    const int tile = tl.tstart[b] + (t0 * tl.nt[b][1] + t1) * tl.nt[b][2] + t2;
    const int slot = ((u0 - t0 * tl.te[b][0]) * tl.te[b][1] + (u1 - t1 * tl.te[b][1])) * kRow + (u2 - t2 * tl.te[b][2]);
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const size_t gi = gidx(a, B.o0[0] + l[0], B.o0[1] + l[1], B.o0[2] + l[2]);
#pragma unroll
    for (int o = 0; o < kOut; ++o) a.Eout[o * N + gi] += (T)ybuf[((size_t)tile * kOut + o) * kTileRows + slot];
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
            const int ext = S.box[b].o1[perm[q]] - S.box[b].o0[perm[q]];
            S.tiles.perm[b][q] = perm[q];
            S.tiles.te[b][q] = te[q];
            S.tiles.nt[b][q] = (ext + te[q] - 1) / te[q];
        }
        S.tiles.tstart[b] = ts;
        ts += S.tiles.nt[b][0] * S.tiles.nt[b][1] * S.tiles.nt[b][2];
    }
    S.tiles.tstart[nb] = ts;
    S.ntiles = ts;
    S.scratch_floats = std::max((size_t)S.hcells * kHid, (size_t)ts * kOut * kTileRows);
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

// Scratch layout: [0, scratch_floats) the correction, tile-major
// [tile][3][kTileRows] FP32 between the two forward launches (the backward
// keeps hidden values there), then the packed weight fragments.
uint16_t* packed_weights(const ShellPlan& S, float* scratch) {
    return reinterpret_cast<uint16_t*>(scratch + S.scratch_floats);
}

// workgroups of shell_mfma resident on the current device at once
int shell_slots(bool f32) {
    static int cached[2][64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int& c = cached[f32 ? 0 : 1][dev];
    if (c == 0) {
        int cu = 0, per_cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) cu = 256;
        const hipError_t e = f32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, shell_mfma<float>, 256, 0)
                                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, shell_mfma<double>, 256, 0);
        if (e != hipSuccess || per_cu <= 0) per_cu = 1;
        c = cu * per_cu;
    }
    return c;
}

void launch_shell_pack(const ShellPlan& S, const uint16_t* w, float* scratch, hipStream_t st) {
    hipLaunchKernelGGL(shell_pack, dim3((S.nbox * kPackBox + 255) / 256), dim3(256), 0, st, w, S.nbox,
                       packed_weights(S, scratch));
}

// The correction kernels on stream st (weights already packed in scratch).
void launch_shell_apply(const Plan& P, const ShellPlan& S, const uint16_t* w, const void* E, const void* H,
                        float* scratch, hipStream_t st) {
    const int g2 = (S.out_cells + 255) / 256;
    const uint16_t* pk = packed_weights(S, scratch);
#ifndef FDTD_SHELL_RUN
#define FDTD_SHELL_RUN 0
#endif
    // FDTD_SHELL_RUN = 0: one tile per workgroup; otherwise runs of
    // ceil(tiles / resident workgroups) tiles
    int per_block = 1;
    if (FDTD_SHELL_RUN) {
        const int slots = shell_slots(P.d.dtype == FDTD_F32);
        per_block = (S.ntiles + slots - 1) / slots;
    }
    const int grid = (S.ntiles + per_block - 1) / per_block;
    if (P.d.dtype == FDTD_F32) {
        ShellArgs<float> a;
        fill_shell<float>(P, S, a);
        a.w = w; a.E = (const float*)E; a.H = (const float*)H; a.Eout = (float*)E;
        hipLaunchKernelGGL(shell_mfma<float>, dim3(grid), dim3(256), 0, st, a, S.tiles, pk, scratch, per_block);
        hipLaunchKernelGGL(shell_add<float>, dim3(g2), dim3(256), 0, st, a, S.tiles, (const float*)scratch, S.out_cells);
    } else {
        ShellArgs<double> a;
        fill_shell<double>(P, S, a);
        a.w = w; a.E = (const double*)E; a.H = (const double*)H; a.Eout = (double*)E;
        hipLaunchKernelGGL(shell_mfma<double>, dim3(grid), dim3(256), 0, st, a, S.tiles, pk, scratch, per_block);
        hipLaunchKernelGGL(shell_add<double>, dim3(g2), dim3(256), 0, st, a, S.tiles, (const float*)scratch, S.out_cells);
    }
}

}  // namespace

extern "C" {

int64_t fdtd_shell_hidden_elems(const void* plan, int32_t thickness) {
    if (!plan) return -1;
    const Plan& P = *(const Plan*)plan;
    if (check_shell(P, thickness)) return -1;
    return (int64_t)make_shell(P, thickness).scratch_floats + kPackFloats;
}

int64_t fdtd_shell_weight_elems(void) { return 6 * (int64_t)(kW1 + kW2); }

int fdtd_shell_apply(const void* plan, int32_t thickness, const uint16_t* w, const void* E, const void* H,
                     float* hidden, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_shell(P, thickness)) return rc;
    if (!w || !E || !H || !hidden) return fail(2, "w, E, H and hidden are required");
    const ShellPlan S = make_shell(P, thickness);
    launch_shell_pack(S, w, hidden, (hipStream_t)stream);
    launch_shell_apply(P, S, w, E, H, hidden, (hipStream_t)stream);
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
    if (int rc = check_fields(P, f, true)) return rc;
    const ShellPlan S = make_shell(P, thickness);
    hipStream_t st = (hipStream_t)stream;
    fdtd_fields cur = *f;
    const size_t esz = P.d.dtype == FDTD_F32 ? 4 : 8;
    const bool swap = (f->H_out && f->H_out != f->H);
    if (n > 0) launch_shell_pack(S, w, hidden, st);
    for (int k = 0; k < n; ++k) {
        const void* amp = src_wave ? (const char*)src_wave + (size_t)k * esz : nullptr;
        int rc = (P.d.dtype == FDTD_F32) ? launch_e<float>(P, &cur, st, (const float*)amp, src_index)
                                         : launch_e<double>(P, &cur, st, (const double*)amp, src_index);
        if (rc) return rc;
        launch_shell_apply(P, S, w, cur.E, cur.H, hidden, st);
        rc = (P.d.dtype == FDTD_F32) ? launch_h<float>(P, &cur, st) : launch_h<double>(P, &cur, st);
        if (rc) return rc;
        if (swap) {
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
            if (P.d.bloch_mask && cur.G_out && cur.G_out != cur.G) {
                std::swap(cur.G, cur.G_out);
                std::swap(cur.psi_g, cur.psi_g_out);
            }
        }
    }
    return check_launch("fdtd_shell_step_n");
}

int64_t fdtd_shell_partials_elems(void) { return (int64_t)64 * kW1; }

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
            hipLaunchKernelGGL(shell_wgrad_sum, dim3((kW2 + 255) / 256), dim3(256), 0, st, (const float*)partials,
                               groups, kW2, wgrad + (size_t)S.nbox * kW1 + (size_t)b * kW2);
        }
        return 0;
    };
    if (P.d.dtype == FDTD_F32) run(0.f);
    else run(0.0);
    return check_launch("fdtd_shell_backward");
}

}  // extern "C"
