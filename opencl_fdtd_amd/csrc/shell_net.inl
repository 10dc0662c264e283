// The learned absorbing shell (part of the fdtd_hip.hip translation unit).
//
// A thin shell of T cells on every face runs the CPML-T update; after the E
// update (and source) of every corrected step (every step, or every second
// or fourth: the caller's period) a small convolutional net adds a
// correction to E on the shell cells:
//   x = (Ex, Ey, Ez, Hx, Hy, Hz)                         6 input channels
//   h_k(j) = sum_{c,t} W1[k][c][t] x_c(j + t)            16 hidden channels
//   y_a(i) = sum_{k,t} W2[d(i)][a][k][t] h_k(i + t)      3 outputs
//   E_a(i) += y_a(i)  for i in the shell
// with d(i) the depth of i: its index along its box's thin axis (a layer-2
// head per depth).
// t runs over the 27 offsets of a 3x3x3 stencil (periodic wrap, as the step).
// There is no activation: the correction must be linear in the fields, as
// Maxwell's equations are.  Weights are BF16.  The inference evaluates the
// two layers as their composed 5x5x5 kernel (composed in FP32, rounded once
// to BF16) on inputs rounded to BF16, with exact products accumulated in
// FP32; the backward (training) works in FP32 on the two layers.
//
// The shell is split into 6 boxes, one per face (x faces own the edges and
// corners they touch, then y faces), each with its own weights.  For the
// backward, hidden values live per box on the box grown by one cell, in a
// caller-provided FP32 scratch [cell][16].

#ifndef FDTD_SHELL_WAVES
#define FDTD_SHELL_WAVES 3           // shell_conv: minimum waves per SIMD (3 workgroups of 4 waves per CU)
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


// wgrad[b][w] += sum_g partials[g][w], fixed order
__global__ __launch_bounds__(256) void shell_wgrad_sum(const float* part, int groups, int nw, float* out) {
    const int wi = blockIdx.x * 256 + threadIdx.x;
    if (wi >= nw) return;
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += (double)part[(size_t)g * nw + wi];
    out[wi] += (float)s;
}

// ---- inference: the composed 5x5x5 convolution on the matrix cores -------
//
// The net has no activation, so per box b and depth d its two layers are one
// 5x5x5 convolution from the 6 inputs to the 3 outputs:
//   Wc[b][d][o][c][s] = sum_k sum_u W2[b][d][o][k][u] W1[b][k][c][s - u]
// over the 3x3x3 offsets u with s - u also one (shell_compose: k outer, u
// inner, FP32 products of BF16 values, which are exact, and FP32 sums, then
// one rounding to BF16), and
//   y_o(i) = sum_{c,s} Wc[b][d(i)][o][c][s] x_c(i + s),   x rounded to BF16.
// That is 2250 multiply-adds per output cell against 3888 for the two layers
// (whose hidden layer a tile must also evaluate on a halo).
//
// One workgroup (4 waves) per tile of a box.  A tile has three axes: a0 the
// box's thin axis (all T depths), a1 (16 rows) and a2 (16 columns, innermost
// in LDS); a2 is z for the x and y faces and y for the z faces.  Its input,
// planes -2 .. 5 along a0 (8 whatever T; weights of planes past T+1 are
// zero) by 20 x 20 cells, is staged in LDS as BF16 [cell][8] (6 channels
// and 2 zero pads), with the a0 stride padded by one cell so that writes in
// grid order (z = a0 on the z faces) meet no bank conflicts.
//
// The convolution is a GEMM with the weights as the A operand (M = 16 rows
// m = 4 d + o: depth d, output o < 3) and 16 cells of one a2 row as B (N).
// K runs over (input plane j, row offset e1, column offset e2, channel):
// a fragment (K = 32) holds four (j, e2) pairs p = 4 q + g (lane group g;
// j = p / 5, e2 = p % 5) of one row offset e1, 8 channels each, and the
// weight of depth d is the one of offset j - 2 - d along a0 (zero beyond
// +-2).  So every MFMA yields all T depths of 16 cells, and the B fragment
// of input row r, group q feeds the MFMAs of output rows r - e1 for the
// five e1: a wave walking its rows reads each B fragment once for up to 5
// MFMAs.  50 fragments (10 q x 5 e1) per box: waves 2 h + k take output
// rows 8 h .. 8 h + 7 and the pairs q = 5 k .. 5 k + 4 (25 fragments,
// 100 VGPRs, held for the whole tile); the two K halves of a row are summed
// through LDS at the end.
// MFMA operands (16x16x32): lane l holds A row l%16 and B column l%16 at
// k = 8*(l/16) + [0..7]; D column l%16 at rows 4*(l/16) + [0..3], i.e. for
// cell l%16 the outputs 0..2 of depth l/16.  The correction goes to a
// scratch laid out as the fields, [3][N] (shell cells only), and is added
// to E later (by the next E update, or by shell_add), since neighbouring
// tiles read E around their own cells.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxT = 4;
constexpr int kCR = 8;                                 // output rows per wave
constexpr int kTileA1 = 2 * kCR, kTileA2 = 16;         // output rows (a1) and columns (a2) per tile
constexpr int kInR = kTileA1 + 4, kInC = kTileA2 + 4;  // input rows and columns
constexpr int kInPlanes = kMaxT + 4;                   // input planes along a0
constexpr int kInA0 = kInR * kInC + 1;                 // LDS cells per input plane (padded)
constexpr int kTileLds = kInPlanes * kInA0;            // uint4 words (51.3 KB)
constexpr int kPairs = kInPlanes * 5;                  // (plane, column offset) pairs
constexpr int kQ = kPairs / 4, kQHalf = kQ / 2;        // pair groups, per K half
constexpr int kFragBox = 5 * kQ;                       // A fragments per box
static_assert(kPairs % 8 == 0, "two K halves of whole pair groups");
static_assert(kDepth >= kMaxT && kDepth * 4 <= 16, "one MFMA row group per depth");
static_assert(kTileA1 * 64 <= kTileLds, "the K-half partials fit the input tile's LDS");

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

// Composed weights, rounded to BF16, as A fragments in lane order: per box
// [K half][e1][q - 5 K half][64 lanes][8 channels], so a lane loads each
// fragment with one 16-byte read.  One thread per BF16 value.
constexpr int kPackBox = kFragBox * 64 * 8;                // BF16 values per box
constexpr int kPackFloats = 6 * kPackBox / 2;              // FP32 words of scratch for 6 boxes

__global__ __launch_bounds__(256) void shell_compose(const uint16_t* w, int nbox, int T, ShellTiles tl,
                                                     uint16_t* packed) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= nbox * kPackBox) return;
    const int b = idx / kPackBox, r = idx % kPackBox;
    const int f = r / 512, lane = (r / 8) % 64, c = r % 8;
    const int kh = f / (5 * kQHalf), e1 = (f / kQHalf) % 5, q = kh * kQHalf + f % kQHalf;
    const int m = lane & 15, g = lane >> 4;
    const int d = m >> 2, o = m & 3;
    const int p = 4 * q + g;
    const int d0 = p / 5 - 2 - d;                          // offset along a0 of plane p / 5 from depth d
    float acc = 0.f;
    if (o < kOut && d < T && c < kIn && d0 >= -2 && d0 <= 2) {
        int s[3];                                          // offset along the grid axes
        const int* pm = tl.perm[b];
        s[pm[0]] = d0; s[pm[1]] = e1 - 2; s[pm[2]] = p % 5 - 2;
        // layer-2 taps u with s - u a layer-1 tap, in increasing tap order
        int lo[3], hi[3];
        for (int q2 = 0; q2 < 3; ++q2) { lo[q2] = max(-1, s[q2] - 1); hi[q2] = min(1, s[q2] + 1); }
        const uint16_t* w1 = w + (size_t)b * kW1 + (size_t)c * kTaps;
        const uint16_t* w2 = w + (size_t)nbox * kW1 + ((size_t)b * kDepth + d) * kW2 + (size_t)o * kHid * kTaps;
        for (int k = 0; k < kHid; ++k) {
            for (int u0 = lo[0]; u0 <= hi[0]; ++u0)
                for (int u1 = lo[1]; u1 <= hi[1]; ++u1)
                    for (int u2 = lo[2]; u2 <= hi[2]; ++u2) {
                        const int u = (u0 + 1) * 9 + (u1 + 1) * 3 + (u2 + 1);
                        const int v = (s[0] - u0 + 1) * 9 + (s[1] - u1 + 1) * 3 + (s[2] - u2 + 1);
                        const float prod = bf16_to_f(w2[k * kTaps + u]) * bf16_to_f(w1[k * kIn * kTaps + v]);
                        acc = acc + prod;
                    }
        }
    }
    packed[idx] = __builtin_bit_cast(uint16_t, (__bf16)acc);
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

template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FDTD_SHELL_WAVES))) void shell_conv(ShellArgs<T> a, ShellTiles tl, const uint16_t* packed, float* ybuf) {
    __shared__ uint4 s_buf[kTileLds];         // the input tile [cell][8 x BF16], then the K-half partials
    // hardware workgroup k runs on XCD k % 8; each XCD gets a contiguous
    // range of tiles so the halos of neighbouring tiles meet in its L2
    const int nb = gridDim.x, qq = nb >> 3, rr = nb & 7, xcd = blockIdx.x & 7, jj = blockIdx.x >> 3;
    const int tile = (xcd < rr ? xcd * (qq + 1) : rr * (qq + 1) + (xcd - rr) * qq) + jj;
    // tiles alternate between the two boxes of an axis (same tile grid), so
    // the two sides of a tile column run together on one XCD: the lines
    // they share (the wrap of a row, or its two ends on the z faces) are
    // fetched from HBM once
    int ax = 0;
    while (ax < 2 && tile >= tl.tstart[2 * ax + 2]) ++ax;
    const int b = 2 * ax + ((tile - tl.tstart[2 * ax]) & 1);
    const ShellBox& B = a.box[b];
    const int p0 = tl.perm[b][0], p1 = tl.perm[b][1], p2 = tl.perm[b][2];
    const int tix = (tile - tl.tstart[2 * ax]) >> 1;
    const int t2 = tix % tl.nt[b][2], t1 = tix / tl.nt[b][2];
    int o[3];                                  // tile origin, grid axes
    o[p0] = B.o0[p0];
    o[p1] = B.o0[p1] + t1 * kTileA1;
    o[p2] = B.o0[p2] + t2 * kTileA2;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int kh = wave & 1, rh = wave >> 1;
    const int g = lane >> 4, n = lane & 15;
    // 0. this wave's 25 A fragments (loaded first: they need nothing else)
    bf16x8 A[5 * kQHalf];
    {
        const uint4* fr = reinterpret_cast<const uint4*>(packed) + ((size_t)b * 2 + kh) * (5 * kQHalf) * 64 + lane;
#pragma unroll
        for (int f = 0; f < 5 * kQHalf; ++f) A[f] = __builtin_bit_cast(bf16x8, fr[f * 64]);
    }
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    // 1. input tile -> LDS (BF16, 8 per cell), loaded in grid order (z
    // fastest) in passes whose loads are all issued before the first
    // conversion (threads past the tile reload its last cell)
    {
        int ext[3], str[3];                    // input extents and LDS strides per grid axis
        ext[p0] = kInPlanes; ext[p1] = kInR; ext[p2] = kInC;
        str[p0] = kInA0; str[p1] = kInC; str[p2] = 1;
        constexpr int nin = kInPlanes * kInR * kInC;
        const float r2 = 1.0f / ext[2], r1 = 1.0f / ext[1];
        const T* E0 = a.E; const T* E1 = a.E + N; const T* E2 = a.E + 2 * N;
        const T* H0 = a.H; const T* H1 = a.H + N; const T* H2 = a.H + 2 * N;
        constexpr int U = sizeof(T) == 4 ? 4 : 2;   // cells per pass
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
    __syncthreads();
    // 2. the GEMM: input row r (of the wave's kCR + 4), pair group q feeds
    // output rows r - e1
    f32x4 acc[kCR];
#pragma unroll
    for (int r = 0; r < kCR; ++r) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
    {
        int off[kQHalf];
#pragma unroll
        for (int k = 0; k < kQHalf; ++k) {
            const int p = 4 * (kh * kQHalf + k) + g;
            off[k] = (p / 5) * kInA0 + rh * kCR * kInC + n + p % 5;
        }
#pragma unroll
        for (int r = 0; r < kCR + 4; ++r) {
#pragma unroll
            for (int k = 0; k < kQHalf; ++k) {
                const uint4 bv = s_buf[off[k] + r * kInC];
#pragma unroll
                for (int e1 = 0; e1 < 5; ++e1) {
                    const int ro = r - e1;
                    if (ro >= 0 && ro < kCR) acc[ro] = mfma(A[e1 * kQHalf + k], bv, acc[ro]);
                }
            }
        }
    }
    // 3. sum the K halves (the odd wave's partials through LDS) and store
    // outputs 0..2 of cell (depth g, row, column n) to the correction, per
    // component on the grid: ycorr[component][grid index]
    __syncthreads();
    f32x4* part = reinterpret_cast<f32x4*>(s_buf) + (size_t)rh * kCR * 64 + lane;
    if (kh == 1) {
#pragma unroll
        for (int r = 0; r < kCR; ++r) part[r * 64] = acc[r];
    }
    __syncthreads();
    if (kh == 1) return;
    int ext[3], gst[3];
    for (int q = 0; q < 3; ++q) ext[q] = B.o1[q] - B.o0[q];
    gst[0] = a.sy * a.sz; gst[1] = a.sz; gst[2] = 1;
    const int l1 = o[p1] - B.o0[p1] + rh * kCR, l2 = o[p2] - B.o0[p2] + n;
    if (g < ext[p0] && l2 < ext[p2]) {
        float* yc = ybuf + (size_t)(B.o0[p0] + g) * gst[p0] + (size_t)(o[p1] + rh * kCR) * gst[p1] +
                    (size_t)(o[p2] + n) * gst[p2];
#pragma unroll
        for (int r = 0; r < kCR; ++r) {
            if (l1 + r < ext[p1]) {
                const f32x4 v = acc[r] + part[r * 64];
                yc[(size_t)r * gst[p1]] = v[0];
                yc[N + (size_t)r * gst[p1]] = v[1];
                yc[2 * N + (size_t)r * gst[p1]] = v[2];
            }
        }
    }
}

// E += correction on the shell cells, one thread per cell in box order
// (z fastest within a box), the correction read at the cell's grid offset.
template <class T>
__global__ __launch_bounds__(256) void shell_add(ShellArgs<T> a, ShellTiles tl, const float* ycorr) {
    int cell = blockIdx.x * 256 + threadIdx.x;
    const int out = tl.ostart[a.nbox];
    if (cell >= out) return;
    int b = 0;
    while (b < a.nbox - 1 && cell >= tl.ostart[b + 1]) ++b;
    cell -= tl.ostart[b];
    const ShellBox& B = a.box[b];
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    const int lz = cell % nz, ly = (cell / nz) % ny, lx = cell / (nz * ny);
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const size_t gi = ((size_t)(B.o0[0] + lx) * a.sy + (B.o0[1] + ly)) * a.sz + (B.o0[2] + lz);
    a.Eout[gi] += (T)ycorr[gi];
    a.Eout[N + gi] += (T)ycorr[N + gi];
    a.Eout[2 * N + gi] += (T)ycorr[2 * N + gi];
}

// ---- training with MFMA (v_mfma_f32_16x16x4_f32) --------------------------
//
// The backward of one application of the net, and the forward hidden layer
// it needs, as GEMMs on the FP32 matrix cores (FP32 operands, so the adjoint
// keeps the precision of FP32 arithmetic; BF16 weights and BF16-rounded
// inputs are exact in FP32).  Operand layout of the 16x16x4 instruction, 64
// lanes: lane l holds A[l%16][l/16] and B[l/16][l%16]; D row 4(l/16)+i,
// column l%16 in its register i.  Every wave works on its own 16 cells (or
// its own column blocks), gathering its B (or A) values from global memory
// (L1/L2-resident neighbourhoods); there is no LDS traffic and no atomics,
// and every sum has a fixed order, so results are bit-reproducible.
//
//   hidden:  h[k](j)   = sum_(c,t)     W1[k][c][t] x_c(j+t)      M=k,    N=16 cells, K=(c,t) 162
//   l2_t:    hbar_k(j) = sum_(d,o,t)   W2[d][o][k][t] ybar_o(i)  M=k,    N=16 cells, K=(d,o,t) 324,
//                        i = j - t, the term kept where i is an output cell of depth d
//   l1_t:    xbar_c(i) = sum_(k,t)     W1[k][c][t] hbar_k(i-t)   M=c(6), N=16 cells, K=(k,t) 432
//   wgrad1:  W1bar[k][(c,t)]   = sum_j hbar_k(j) x_c(j+t)        M=k, N=(c,t) in blocks of 16, K=cells
//   wgrad2:  W2bar[d][o][(k,t)] = sum_i [d(i)=d] ybar_o(i) h_k(i+t)  M=(d,o) 12, N=(k,t) blocks, K=cells

__device__ __forceinline__ f32x4 mfma4(float a, float b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// prefix of 16-cell groups per box (hidden cells)
struct ShellGroups {
    int gstart[7];
};

__device__ __forceinline__ void hid_coords(const ShellBox& B, int l, int* lx, int* ly, int* lz) {
    *lz = l % B.e[2];
    *ly = (l / B.e[2]) % B.e[1];
    *lx = l / (B.e[2] * B.e[1]);
}

// layer 1 over every grown-box cell (the forward hidden values, BF16)
template <class T>
__global__ __launch_bounds__(256) void shell_l1_mfma(ShellArgs<T> a, ShellGroups gr) {
    const int grp = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (grp >= gr.gstart[a.nbox]) return;                  // whole waves
    int b = 0;
    while (b < a.nbox - 1 && grp >= gr.gstart[b + 1]) ++b;
    const ShellBox& B = a.box[b];
    const int lane = threadIdx.x & 63, n = lane & 15, q = lane >> 4;
    const int l = (grp - gr.gstart[b]) * 16 + n;
    const bool ok = l < B.hcells;
    int lx = 0, ly = 0, lz = 0;
    if (ok) hid_coords(B, l, &lx, &ly, &lz);
    const int x = hid_global(B, 0, lx), y = hid_global(B, 1, ly), z = hid_global(B, 2, lz);
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const uint16_t* w1 = a.w + (size_t)b * kW1;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int s4 = 0; s4 < (kIn * kTaps + 3) / 4; ++s4) {
        const int kk = 4 * s4 + q;                          // (c, t) = (kk / 27, kk % 27)
        float av = 0.f, bv = 0.f;
        if (kk < kIn * kTaps) {
            const int c = kk / kTaps, t = kk % kTaps;
            av = bf16_to_f(w1[(n * kIn + c) * kTaps + t]);
            if (ok) {
                const size_t gi = gidx(a, x + t / 9 - 1, y + (t / 3) % 3 - 1, z + t % 3 - 1);
                bv = bf16_round((float)(c < 3 ? a.E[c * N + gi] : a.H[(c - 3) * N + gi]));
            }
        }
        acc = mfma4(av, bv, acc);
    }
    if (ok) {
        float* h = a.hid + ((size_t)B.hoff + l) * kHid + 4 * q;
#pragma unroll
        for (int i = 0; i < 4; ++i) h[i] = bf16_round(acc[i]);
    }
}

// hbar over every grown-box cell from ybar = Ebar on the output cells
template <class T>
__global__ __launch_bounds__(256) void shell_l2_t_mfma(ShellArgs<T> a, ShellGroups gr) {
    const int grp = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (grp >= gr.gstart[a.nbox]) return;
    int b = 0;
    while (b < a.nbox - 1 && grp >= gr.gstart[b + 1]) ++b;
    const ShellBox& B = a.box[b];
    const int lane = threadIdx.x & 63, n = lane & 15, q = lane >> 4;
    const int l = (grp - gr.gstart[b]) * 16 + n;
    const bool ok = l < B.hcells;
    int lx = 0, ly = 0, lz = 0;
    if (ok) hid_coords(B, l, &lx, &ly, &lz);
    const int gx = hid_global(B, 0, lx), gy = hid_global(B, 1, ly), gz = hid_global(B, 2, lz);
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const uint16_t* w2b = a.w + (size_t)a.nbox * kW1 + (size_t)b * kDepth * kW2;
    const int r = b / 2;                                   // the box's thin axis
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int s4 = 0; s4 < kDepth * kOut * kTaps / 4; ++s4) {
        const int kk = 4 * s4 + q;                          // (d, o, t)
        const int t = kk % kTaps, o = (kk / kTaps) % kOut, d = kk / (kTaps * kOut);
        const float av = bf16_to_f(w2b[(size_t)d * kW2 + (o * kHid + n) * kTaps + t]);
        float bv = 0.f;
        if (ok) {
            const int ox = out_local(B, 0, gx - (t / 9 - 1));
            const int oy = out_local(B, 1, gy - ((t / 3) % 3 - 1));
            const int oz = out_local(B, 2, gz - (t % 3 - 1));
            if (ox >= 0 && oy >= 0 && oz >= 0 && (r == 0 ? ox : (r == 1 ? oy : oz)) == d)
                bv = (float)a.Ebar[o * N + gidx(a, B.o0[0] + ox, B.o0[1] + oy, B.o0[2] + oz)];
        }
        acc = mfma4(av, bv, acc);
    }
    if (ok) {
        float* hb = a.hbar + ((size_t)B.hoff + l) * kHid + 4 * q;
#pragma unroll
        for (int i = 0; i < 4; ++i) hb[i] = acc[i];
    }
}

// xbar over the cells of box part_box grown by two, accumulated into
// Ebar_out / Hbar_out (boxes one launch after another: their regions overlap)
template <class T>
__global__ __launch_bounds__(256) void shell_l1_t_mfma(ShellArgs<T> a) {
    const ShellBox& B = a.box[a.part_box];
    const int g0 = in_extent(B, 0), g1 = in_extent(B, 1), g2 = in_extent(B, 2);
    const int ncell = g0 * g1 * g2;
    const int grp = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (grp * 16 >= ncell) return;
    const int lane = threadIdx.x & 63, n = lane & 15, q = lane >> 4;
    const int cell = grp * 16 + n;
    const bool ok = cell < ncell;
    const int lz = ok ? cell % g2 : 0, ly = ok ? (cell / g2) % g1 : 0, lx = ok ? cell / (g2 * g1) : 0;
    const int gx = in_global(B, 0, lx), gy = in_global(B, 1, ly), gz = in_global(B, 2, lz);
    const uint16_t* w1 = a.w + (size_t)a.part_box * kW1;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int s4 = 0; s4 < kHid * kTaps / 4; ++s4) {
        const int kk = 4 * s4 + q;                          // (k, t)
        const int k = kk / kTaps, t = kk % kTaps;
        const float av = n < kIn ? bf16_to_f(w1[(k * kIn + n) * kTaps + t]) : 0.f;
        float bv = 0.f;
        if (ok) {
            const int hx = hid_local(B, 0, gx - (t / 9 - 1));
            const int hy = hid_local(B, 1, gy - ((t / 3) % 3 - 1));
            const int hz = hid_local(B, 2, gz - (t % 3 - 1));
            if (hx >= 0 && hy >= 0 && hz >= 0)
                bv = a.hbar[((size_t)B.hoff + ((size_t)hx * B.e[1] + hy) * B.e[2] + hz) * kHid + k];
        }
        acc = mfma4(av, bv, acc);
    }
    // rows c = 4q + i: E components 0..2 (q = 0), H components 0..2 (c = 3..5)
    if (ok && q < 2) {
        const size_t N = (size_t)a.sx * a.sy * a.sz;
        const size_t gi = gidx(a, gx, gy, gz);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = 4 * q + i;
            if (c < 3) a.Ebar_out[c * N + gi] += (T)acc[i];
            else if (c < kIn) a.Hbar_out[(c - 3) * N + gi] += (T)acc[i];
        }
    }
}

// Weight-gradient partials of box part_box: workgroup g takes a contiguous
// slice of the cells; its waves split the output column blocks and each
// walks the whole slice, 4 cells per MFMA.  partials[g][w], in the weights'
// own layout, summed in a fixed order by shell_wgrad_sum.
constexpr int kWgGroups = 64;

template <class T>
__global__ __launch_bounds__(256) void shell_wgrad1_mfma(ShellArgs<T> a) {
    constexpr int NCOL = kIn * kTaps;                       // 162 columns (c, t)
    constexpr int NBLK = (NCOL + 15) / 16;                  // 11 blocks of 16
    constexpr int PER_WAVE = (NBLK + 3) / 4;                // 3
    const ShellBox& B = a.box[a.part_box];
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const int per = (B.hcells + gridDim.x - 1) / gridDim.x;
    const int c0 = blockIdx.x * per, c1 = min(c0 + per, B.hcells);
    const int lane = threadIdx.x & 63, n = lane & 15, q = lane >> 4;
    const int wave = (int)(threadIdx.x >> 6);
    f32x4 acc[PER_WAVE];
#pragma unroll
    for (int j = 0; j < PER_WAVE; ++j) acc[j] = {0.f, 0.f, 0.f, 0.f};
    // this lane's B columns (c, t) per block
    int cc[PER_WAVE], tt[PER_WAVE];
#pragma unroll
    for (int j = 0; j < PER_WAVE; ++j) {
        const int col = 16 * (wave + 4 * j) + n;
        cc[j] = col < NCOL ? col / kTaps : -1;
        tt[j] = col < NCOL ? col % kTaps : 0;
    }
    for (int base = c0; base < c1; base += 4) {
        const int l = base + q;                             // this lane's cell (K index)
        float av = 0.f;
        int x = 0, y = 0, z = 0;
        const bool ok = l < c1;
        if (ok) {
            av = a.hbar[((size_t)B.hoff + l) * kHid + n];   // A[k = n][cell q]
            int lx, ly, lz;
            hid_coords(B, l, &lx, &ly, &lz);
            x = hid_global(B, 0, lx); y = hid_global(B, 1, ly); z = hid_global(B, 2, lz);
        }
#pragma unroll
        for (int j = 0; j < PER_WAVE; ++j) {
            if (wave + 4 * j >= NBLK) break;                // uniform per wave
            float bv = 0.f;
            if (ok && cc[j] >= 0) {
                const int t = tt[j], c = cc[j];
                const size_t gi = gidx(a, x + t / 9 - 1, y + (t / 3) % 3 - 1, z + t % 3 - 1);
                bv = bf16_round((float)(c < 3 ? a.E[c * N + gi] : a.H[(c - 3) * N + gi]));
            }
            acc[j] = mfma4(av, bv, acc[j]);
        }
    }
    float* out = a.wgrad_part + (size_t)blockIdx.x * kW1;
#pragma unroll
    for (int j = 0; j < PER_WAVE; ++j) {
        const int col = 16 * (wave + 4 * j) + n;
        if (wave + 4 * j < NBLK && col < NCOL) {
#pragma unroll
            for (int i = 0; i < 4; ++i) out[(4 * q + i) * NCOL + col] = acc[j][i];   // W1[k][(c, t)]
        }
    }
}

template <class T>
__global__ __launch_bounds__(256) void shell_wgrad2_mfma(ShellArgs<T> a) {
    constexpr int NCOL = kHid * kTaps;                      // 432 columns (k, t)
    constexpr int NBLK = NCOL / 16;                         // 27
    constexpr int PER_WAVE = (NBLK + 3) / 4;                // 7
    static_assert(kDepth * kOut <= 16, "rows (depth, output) fit one MFMA");
    const ShellBox& B = a.box[a.part_box];
    const size_t N = (size_t)a.sx * a.sy * a.sz;
    const int ny = B.o1[1] - B.o0[1], nz = B.o1[2] - B.o0[2];
    const int nout = (B.o1[0] - B.o0[0]) * ny * nz;
    const int per = (nout + gridDim.x - 1) / gridDim.x;
    const int c0 = blockIdx.x * per, c1 = min(c0 + per, nout);
    const int r = a.part_box / 2;
    const int lane = threadIdx.x & 63, n = lane & 15, q = lane >> 4;
    const int wave = (int)(threadIdx.x >> 6);
    const int ad = n / kOut, ao = n % kOut;                  // this lane's A row (depth, output)
    f32x4 acc[PER_WAVE];
#pragma unroll
    for (int j = 0; j < PER_WAVE; ++j) acc[j] = {0.f, 0.f, 0.f, 0.f};
    int kk_[PER_WAVE], tt[PER_WAVE];
#pragma unroll
    for (int j = 0; j < PER_WAVE; ++j) {
        const int col = 16 * (wave + 4 * j) + n;
        kk_[j] = col / kTaps;
        tt[j] = col % kTaps;
    }
    for (int base = c0; base < c1; base += 4) {
        const int l = base + q;
        const bool ok = l < c1;
        float av = 0.f;
        int ox = 0, oy = 0, oz = 0;
        if (ok) {
            oz = l % nz; oy = (l / nz) % ny; ox = l / (nz * ny);
            const int dep = r == 0 ? ox : (r == 1 ? oy : oz);
            if (n < kDepth * kOut && dep == ad)
                av = (float)a.Ebar[ao * N + gidx(a, B.o0[0] + ox, B.o0[1] + oy, B.o0[2] + oz)];
        }
#pragma unroll
        for (int j = 0; j < PER_WAVE; ++j) {
            if (wave + 4 * j >= NBLK) break;                // uniform per wave
            float bv = 0.f;
            if (ok) {
                const int t = tt[j];
                const int hx = hid_local(B, 0, B.o0[0] + ox + t / 9 - 1);
                const int hy = hid_local(B, 1, B.o0[1] + oy + (t / 3) % 3 - 1);
                const int hz = hid_local(B, 2, B.o0[2] + oz + t % 3 - 1);
                bv = a.hid[((size_t)B.hoff + ((size_t)hx * B.e[1] + hy) * B.e[2] + hz) * kHid + kk_[j]];
            }
            acc[j] = mfma4(av, bv, acc[j]);
        }
    }
    float* out = a.wgrad_part + (size_t)blockIdx.x * (kDepth * kW2);
#pragma unroll
    for (int j = 0; j < PER_WAVE; ++j) {
        const int col = 16 * (wave + 4 * j) + n;
        if (wave + 4 * j < NBLK) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 4 * q + i;                     // row (depth, output)
                if (m < kDepth * kOut) out[(size_t)(m / kOut) * kW2 + (m % kOut) * NCOL + col] = acc[j][i];
            }
        }
    }
}

#ifdef FDTD_RECORD_LAUNCHES
// launch recording of the shell's kernels: the fields the correction reads
// (E_IN, H_IN) and the E it is added to (E_OUT)
template <class T>
void rec_fill(fdtd_launch_record& r, const ShellArgs<T>& a) {
    r.src_comp = -1;
    r.buf[FDTD_REC_E_IN] = a.E;
    r.buf[FDTD_REC_H_IN] = a.H;
    r.buf[FDTD_REC_E_OUT] = a.Eout;
}
#endif

struct ShellPlan {
    int nbox, T;
    ShellBox box[6];
    int hcells, out_cells;
    ShellTiles tiles;
    int ntiles;
    size_t scratch_floats;   // hidden values (backward) or tile-major correction (forward)
};

ShellPlan make_shell(const Plan& P, int T) {
    ShellPlan S;
    S.T = T;
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
    // for z faces (16), a1 = the remaining axis (16)
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
    S.scratch_floats = std::max((size_t)S.hcells * kHid, (size_t)3 * P.N);
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

// Scratch layout: [0, scratch_floats) the correction, [3][N] FP32 on the
// grid (written and read on the shell cells only), between its computation
// and its addition to E (the backward keeps hidden values there), then the
// packed weight fragments.
uint16_t* packed_weights(const ShellPlan& S, float* scratch) {
    return reinterpret_cast<uint16_t*>(scratch + S.scratch_floats);
}

void launch_shell_pack(const ShellPlan& S, const uint16_t* w, float* scratch, hipStream_t st) {
    hipLaunchKernelGGL(shell_compose, dim3((S.nbox * kPackBox + 255) / 256), dim3(256), 0, st, w, S.nbox, S.T,
                       S.tiles, packed_weights(S, scratch));
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

void launch_shell_conv(const Plan& P, const ShellPlan& S, const uint16_t* w, const void* E, const void* H,
                       float* scratch, hipStream_t st) {
    const uint16_t* pk = packed_weights(S, scratch);
    if (P.d.dtype == FDTD_F32)
        hipLaunchKernelGGL(shell_conv<float>, dim3(S.ntiles), dim3(256), 0, st, shell_args<float>(P, S, w, E, H),
                           S.tiles, pk, scratch);
    else
        hipLaunchKernelGGL(shell_conv<double>, dim3(S.ntiles), dim3(256), 0, st, shell_args<double>(P, S, w, E, H),
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
    launch_shell_conv(P, S, w, E, H, hidden, st);
    launch_shell_add(P, S, E, hidden, st);
    return check_launch("fdtd_shell_apply");
}

int fdtd_shell_step_n(const void* plan, const fdtd_fields* f, int32_t thickness, const uint16_t* w, float* hidden,
                      int32_t n, int32_t period, int64_t src_index, const void* src_wave, void* stream,
                      int32_t* moved) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (int rc = check_shell(P, thickness)) return rc;
    if (!w || !hidden) return fail(2, "w and hidden are required");
    if (n < 0) return fail(3, "n = %d < 0", n);
    if (period != 1 && period != 2 && period != 4) return fail(3, "correction period %d is not 1, 2 or 4", period);
    if (n % period) return fail(3, "n = %d is not a multiple of the correction period %d", n, period);
    if (src_wave && (src_index < 0 || src_index >= 3 * P.N))
        return fail(3, "src_index %lld out of [0, %lld)", (long long)src_index, (long long)(3 * P.N));
    if (P.d.poynting == FDTD_POYNTING_TWO_PASS) return fail(3, "two-pass Poynting is not supported with the shell");
    if (P.d.bloch_mask) return fail(3, "the shell correction is for real fields (no Bloch boundaries)");
    if (int rc = check_fields(P, f, true)) return rc;
    if (moved) *moved = 0;
    const ShellPlan S = make_shell(P, thickness);
    hipStream_t st = (hipStream_t)stream;
    fdtd_fields cur = *f;
    const size_t esz = P.d.dtype == FDTD_F32 ? 4 : 8;
    const bool swap = (f->H_out && f->H_out != f->H);
    if (n > 0) launch_shell_pack(S, w, hidden, st);
    const auto amp_of = [&](int k) -> const void* {
        return src_wave ? (const char*)src_wave + (size_t)k * esz : nullptr;
    };
    const auto half_e = [&](int k, bool pending) {
        const float* y = pending ? hidden : nullptr;
        return (P.d.dtype == FDTD_F32)
                   ? launch_e<float>(P, &cur, st, (const float*)amp_of(k), src_index, y, thickness, P.N)
                   : launch_e<double>(P, &cur, st, (const double*)amp_of(k), src_index, y, thickness, P.N);
    };
    const auto half_h = [&]() {
        const int rc = (P.d.dtype == FDTD_F32) ? launch_h<float>(P, &cur, st) : launch_h<double>(P, &cur, st);
        if (swap) {
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
        }
        return rc;
    };
    const auto done = [&]() {
        if (moved)
            *moved = (cur.E != f->E ? 1 : 0) | (cur.psi_e != f->psi_e ? 2 : 0) | (cur.H != f->H ? 4 : 0);
        return check_launch("fdtd_shell_step_n");
    };
    if (period > 1) {
        // step k: E update (+ the correction of step k-1 when k-1 is a
        // multiple of the period) and source; the correction y(E, H) into
        // the scratch when k is a multiple of the period; H update from the
        // uncorrected E.  With the two-step pass (E may end in E_out,
        // reported through *moved): E_0 and its correction, then per pair
        // of steps k, k+1 one pass (H_{k-1}, E_k with the pending
        // correction added to E^{k-1} when there is one, H_k, E_{k+1}) and
        // the correction of E_{k+1} when it is due; the last pass stops
        // after H_{n-1}.
        if (n == 0) return done();
        const bool pass2 = moved && P.d.dtype == FDTD_F32 && !P.d.nonuniform && he_buffers(P, f) &&
                           he_supported(P) && he2_supported(P) &&
                           (P.d.n_face == 0 || (f->psi_e_out && f->psi_e_out != f->psi_e));
        if (pass2) {
            if (int rc = half_e(0, false)) return rc;
            launch_shell_conv(P, S, w, cur.E, cur.H, hidden, st);
            for (int k = 1; k < n; k += 2) {
                const bool last_e = k + 1 < n;
                const bool pending = (k - 1) % period == 0;
                if (int rc = launch_he2(P, &cur, st, (const float*)amp_of(k), src_index, last_e,
                                        pending ? hidden : nullptr, thickness, P.N))
                    return rc;
                std::swap(cur.H, cur.H_out);
                std::swap(cur.psi_h, cur.psi_h_out);
                std::swap(cur.E, cur.E_out);
                std::swap(cur.psi_e, cur.psi_e_out);
                if (last_e && (k + 1) % period == 0) launch_shell_conv(P, S, w, cur.E, cur.H, hidden, st);
            }
            return done();
        }
        for (int k = 0; k < n; ++k) {
            if (int rc = half_e(k, k > 0 && (k - 1) % period == 0)) return rc;
            if (k % period == 0) launch_shell_conv(P, S, w, cur.E, cur.H, hidden, st);
            if (int rc = half_h()) return rc;
        }
        return done();
    }
    // period 1: step k: E update (+ the correction of step k-1, added
    // before the update) and source; correction y(E, H) into the scratch; H
    // update from the uncorrected E.  The last step's correction is added at
    // the end.  With a second E buffer and a plan that has the fused pass:
    // E_0 and its correction, then per step one pass for H_{k-1} + S and E_k
    // (+ the correction of step k-1) and the correction of step k, then
    // H_{n-1}.  Same order of operations per cell; the count of passes is
    // kept even so that E ends in f->E.
    if (n >= 2 && he_buffers(P, f) && he_supported(P)) {
        if (int rc = half_e(0, false)) return rc;
        launch_shell_conv(P, S, w, cur.E, cur.H, hidden, st);
        int k = 1;
        if ((n - 1) % 2 == 1) {
            if (int rc = half_h()) return rc;
            if (int rc = half_e(1, true)) return rc;
            launch_shell_conv(P, S, w, cur.E, cur.H, hidden, st);
            k = 2;
        }
        for (; k < n; ++k) {
            const int rc = (P.d.dtype == FDTD_F32)
                               ? launch_he<float>(P, &cur, st, (const float*)amp_of(k), src_index, hidden,
                                                  thickness, P.N)
                               : launch_he<double>(P, &cur, st, (const double*)amp_of(k), src_index, hidden,
                                                   thickness, P.N);
            if (rc) return rc;
            std::swap(cur.H, cur.H_out);
            std::swap(cur.psi_h, cur.psi_h_out);
            std::swap(cur.E, cur.E_out);
            launch_shell_conv(P, S, w, cur.E, cur.H, hidden, st);
        }
        if (int rc = half_h()) return rc;
        launch_shell_add(P, S, cur.E, hidden, st);
        return done();
    }
    for (int k = 0; k < n; ++k) {
        if (int rc = half_e(k, k > 0)) return rc;
        launch_shell_conv(P, S, w, cur.E, cur.H, hidden, st);
        if (int rc = half_h()) return rc;
    }
    if (n > 0) launch_shell_add(P, S, cur.E, hidden, st);
    return done();
}

int64_t fdtd_shell_partials_elems(void) { return (int64_t)kWgGroups * std::max(kW1, kDepth * kW2); }

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
    const int groups = kWgGroups;
    auto run = [&](auto tag) -> int {
        using T = decltype(tag);
        ShellArgs<T> a;
        fill_shell<T>(P, S, a);
        a.w = w; a.E = (const T*)E; a.H = (const T*)H; a.hid = hidden; a.hbar = hbar;
        a.Ebar = (const T*)Ebar; a.Ebar_out = (T*)Ebar_out; a.Hbar_out = (T*)Hbar_out;
        a.wgrad_part = partials;
        // 16-cell groups of hidden cells, box by box (one group per wave)
        ShellGroups gr;
        gr.gstart[0] = 0;
        for (int b = 0; b < 6; ++b)
            gr.gstart[b + 1] = gr.gstart[b] + (b < S.nbox ? (S.box[b].hcells + 15) / 16 : 0);
        const int g1 = (gr.gstart[S.nbox] + 3) / 4;
        // forward hidden (the weight gradient of layer 2 needs it)
        hipLaunchKernelGGL(shell_l1_mfma<T>, dim3(g1), dim3(256), 0, st, a, gr);
        hipLaunchKernelGGL(shell_l2_t_mfma<T>, dim3(g1), dim3(256), 0, st, a, gr);
        for (int b = 0; b < S.nbox; ++b) {
            a.part_box = b;
            const ShellBox& B = S.box[b];
            const int n2 = (B.full[0] ? B.s[0] : B.o1[0] - B.o0[0] + 4) *
                           (B.full[1] ? B.s[1] : B.o1[1] - B.o0[1] + 4) *
                           (B.full[2] ? B.s[2] : B.o1[2] - B.o0[2] + 4);
            hipLaunchKernelGGL(shell_l1_t_mfma<T>, dim3((n2 + 63) / 64), dim3(256), 0, st, a);
            // W1 gradient of box b
            hipLaunchKernelGGL(shell_wgrad1_mfma<T>, dim3(groups), dim3(256), 0, st, a);
            hipLaunchKernelGGL(shell_wgrad_sum, dim3((kW1 + 255) / 256), dim3(256), 0, st, (const float*)partials,
                               groups, kW1, wgrad + (size_t)b * kW1);
            // W2 gradient of box b
            hipLaunchKernelGGL(shell_wgrad2_mfma<T>, dim3(groups), dim3(256), 0, st, a);
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
