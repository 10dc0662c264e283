// Discrete adjoint of one FDTD step and FP64 reductions (part of the
// fdtd_hip.hip translation unit).
//
// Both half-steps have the form, per cell i and component a,
//   F'_a = F_a + kappa_a * sum_b sigma_ab * (g_cb + psi'_ab)
//   g_cb = p2_b * inv_b * D_b G_c,   psi'_ab = p0_b psi_ab + p1_b g_cb
// with (kappa, D, G) = (dt/eps, backward difference, H) for the E step and
// (-dt, forward difference, E) for the H step (update_e.cl:39-107,
// update_h.cl:36-131; p0/p1/p2 = 0/0/1 outside the slab of axis b).  The
// signs sigma_ab are (x,y,z:+) (x,z,y:-) (y,z,x:+) (y,x,z:-) (z,x,y:+)
// (z,y,x:-) for (a, b, c).  Reversing it, with Fbar' the adjoint of F':
//   u_ab = kappa_a sigma_ab Fbar'_a,  psibar_tot = u + psibar'_ab (in slab)
//   gbar_cb = u + p1 psibar_tot,      psibar_ab(before) = p0 psibar_tot
//   h_cb = p2 inv_b gbar_cb
//   E step: Hbar_c(j) += h_cb(j) - h_cb(j + b)   (transpose of D_backward)
//   H step: Ebar_c(j) += h_cb(j - b) - h_cb(j)   (transpose of D_forward)
//   deps_a += -Ebar'_a (E'_a - E_a) / eps_a       (d kappa / d eps)
// Every kernel gathers: a thread writes only its own cell, so no atomics and
// bit-reproducible results.

namespace {

template <class T>
struct AdjArgs {
    Geom<T> g;
    const T* Ep[3];
    const T* En[3];
    const T* Hp[3];
    const T* Hn[3];
    const T* eps[3];
    T* Eb[3];
    T* Hb[3];
    const T* pin;        // adjoint psi of this half-step, after it
    T* pout;             // ... before it
    double* grad[3];
    int src_x, src_p, src_comp;
    T src_amp;
    int obj_axis, obj_index, obj_mode;
    int glo[3], ghi[3];  // the box of cells whose eps gradient is accumulated
};

struct SlabHit {
    bool in;
    int ir, slot;
};

template <class T>
__device__ __forceinline__ SlabHit find_slab(const Geom<T>& g, int b, int c) {
    SlabHit h{false, 0, 0};
    const int s = b == 0 ? g.sx : (b == 1 ? g.sy : g.sz);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const FaceDev<T>& f = g.face[b][k];
        if (f.thick && c >= f.lo && c < f.hi) {
            h.in = true;
            h.ir = (f.pol == 0) ? c : (s - 1) - c;
            h.slot = k;
        }
    }
    return h;
}

__device__ __forceinline__ int wrap(int c, int s) { return c < 0 ? c + s : (c >= s ? c - s : c); }

// h_cb of pair (A, B) at cell (x, y, z); writes the adjoint psi of the cell
// when write is set.  kappa = dt/eps_A (E step) or -dt (H step).
template <class T, int A, int B, bool ISE>
__device__ __forceinline__ T adj_term(const AdjArgs<T>& q, int x, int y, int z, T kappa, T fbar, bool write) {
    constexpr int C = 3 - A - B;
    constexpr T sigma = ((A == 0 && B == 1) || (A == 1 && B == 2) || (A == 2 && B == 0)) ? (T)1 : (T)-1;
    (void)C;
    const Geom<T>& g = q.g;
    const T u = kappa * sigma * fbar;
    const int c = B == 0 ? x : (B == 1 ? y : z);
    const SlabHit hit = find_slab<T>(g, B, c);
    if (!hit.in) return g.inv_u * u;
    const FaceDev<T>& f = g.face[B][hit.slot];
    const T* cf = g.coef + f.coef_off + (ISE ? 0 : 3 * f.thick);
    const T p0 = cf[hit.ir], p1 = cf[f.thick + hit.ir], p2 = cf[2 * f.thick + hit.ir];
    // component A is psi_u of face B when A is the lower transverse axis
    constexpr int U = B == 0 ? 1 : 0;
    const int p = y * g.sz + z;
    const uint32_t ip = (uint32_t)f.psi_off + (A == U ? 0u : (uint32_t)f.face_elems) +
                        psi_index<B>(g, f, x, y, z, p, hit.ir);
    const T tot = u + q.pin[ip];
    if (write) q.pout[ip] = p0 * tot;
    return p2 * g.inv_u * (u + p1 * tot);
}

template <class T>
__device__ __forceinline__ T fld(const T* const* F, int a, const Geom<T>& g, int x, int y, int z) {
    return F[a][((size_t)x * g.sy + y) * g.sz + z];
}

// E-step adjoint: Hbar += transposed curl of Ebar; adjoint psi_E; deps.
template <class T>
__global__ __launch_bounds__(kBlock) void adj_e(AdjArgs<T> q) {
    const Geom<T>& g = q.g;
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= g.plane) return;
    const int x = blockIdx.y;
    const int y = p / g.sz, z = p - (p / g.sz) * g.sz;
    const size_t i = (size_t)x * g.plane + p;
    const int xp = wrap(x + 1, g.sx), yp = wrap(y + 1, g.sy), zp = wrap(z + 1, g.sz);
    auto kap = [&](int a, int xx, int yy, int zz) { return g.dt / fld(q.eps, a, g, xx, yy, zz); };
    auto eb = [&](int a, int xx, int yy, int zz) { return fld((const T* const*)q.Eb, a, g, xx, yy, zz); };
    // H_x: (a=y, b=z) and (a=z, b=y)
    const T hx = adj_term<T, 1, 2, true>(q, x, y, z, kap(1, x, y, z), eb(1, x, y, z), true) -
                 adj_term<T, 1, 2, true>(q, x, y, zp, kap(1, x, y, zp), eb(1, x, y, zp), false) +
                 adj_term<T, 2, 1, true>(q, x, y, z, kap(2, x, y, z), eb(2, x, y, z), true) -
                 adj_term<T, 2, 1, true>(q, x, yp, z, kap(2, x, yp, z), eb(2, x, yp, z), false);
    // H_y: (a=x, b=z) and (a=z, b=x)
    const T hy = adj_term<T, 0, 2, true>(q, x, y, z, kap(0, x, y, z), eb(0, x, y, z), true) -
                 adj_term<T, 0, 2, true>(q, x, y, zp, kap(0, x, y, zp), eb(0, x, y, zp), false) +
                 adj_term<T, 2, 0, true>(q, x, y, z, kap(2, x, y, z), eb(2, x, y, z), true) -
                 adj_term<T, 2, 0, true>(q, xp, y, z, kap(2, xp, y, z), eb(2, xp, y, z), false);
    // H_z: (a=x, b=y) and (a=y, b=x)
    const T hz = adj_term<T, 0, 1, true>(q, x, y, z, kap(0, x, y, z), eb(0, x, y, z), true) -
                 adj_term<T, 0, 1, true>(q, x, yp, z, kap(0, x, yp, z), eb(0, x, yp, z), false) +
                 adj_term<T, 1, 0, true>(q, x, y, z, kap(1, x, y, z), eb(1, x, y, z), true) -
                 adj_term<T, 1, 0, true>(q, xp, y, z, kap(1, xp, y, z), eb(1, xp, y, z), false);
    q.Hb[0][i] += hx;
    q.Hb[1][i] += hy;
    q.Hb[2][i] += hz;
    // d kappa_a / d eps_a = -kappa_a / eps_a; the bracket is (E' - E) / kappa
    // (only in the gradient's box: outside it E and the gradient are not read)
    if (!q.grad[0] || x < q.glo[0] || x >= q.ghi[0] || y < q.glo[1] || y >= q.ghi[1] || z < q.glo[2] ||
        z >= q.ghi[2])
        return;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double en = (double)q.En[a][i];
        if (a == q.src_comp && x == q.src_x && p == q.src_p) en -= (double)q.src_amp;
        const double de = en - (double)q.Ep[a][i];
        const double ep = (double)q.eps[a][i];
        if (isfinite(ep)) q.grad[a][i] += -(double)q.Eb[a][i] * de / ep;
    }
}

// H-step adjoint: Ebar += transposed curl of Hbar; adjoint psi_H.
template <class T>
__global__ __launch_bounds__(kBlock) void adj_h(AdjArgs<T> q) {
    const Geom<T>& g = q.g;
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= g.plane) return;
    const int x = blockIdx.y;
    const int y = p / g.sz, z = p - (p / g.sz) * g.sz;
    const size_t i = (size_t)x * g.plane + p;
    const int xm = wrap(x - 1, g.sx), ym = wrap(y - 1, g.sy), zm = wrap(z - 1, g.sz);
    const T k = -g.dt;
    auto hb = [&](int a, int xx, int yy, int zz) { return fld((const T* const*)q.Hb, a, g, xx, yy, zz); };
    // E_x: (a=y, b=z) and (a=z, b=y)
    const T ex = adj_term<T, 1, 2, false>(q, x, y, zm, k, hb(1, x, y, zm), false) -
                 adj_term<T, 1, 2, false>(q, x, y, z, k, hb(1, x, y, z), true) +
                 adj_term<T, 2, 1, false>(q, x, ym, z, k, hb(2, x, ym, z), false) -
                 adj_term<T, 2, 1, false>(q, x, y, z, k, hb(2, x, y, z), true);
    const T ey = adj_term<T, 0, 2, false>(q, x, y, zm, k, hb(0, x, y, zm), false) -
                 adj_term<T, 0, 2, false>(q, x, y, z, k, hb(0, x, y, z), true) +
                 adj_term<T, 2, 0, false>(q, xm, y, z, k, hb(2, xm, y, z), false) -
                 adj_term<T, 2, 0, false>(q, x, y, z, k, hb(2, x, y, z), true);
    const T ez = adj_term<T, 0, 1, false>(q, x, ym, z, k, hb(0, x, ym, z), false) -
                 adj_term<T, 0, 1, false>(q, x, y, z, k, hb(0, x, y, z), true) +
                 adj_term<T, 1, 0, false>(q, xm, y, z, k, hb(1, xm, y, z), false) -
                 adj_term<T, 1, 0, false>(q, x, y, z, k, hb(1, x, y, z), true);
    q.Eb[0][i] += ex;
    q.Eb[1][i] += ey;
    q.Eb[2][i] += ez;
}

// Objective J = sum_plane S_d (update_s.cl:33-36 summed over the plane):
// every oS product of component d on the plane enters twice with dt/8, so
// dJ/d(oS) = dt/4.  d = x: +aEyx*aHz, -aEzx*aHy; d = y: +aEzy*aHx,
// -aExy*aHz; d = z: +aExz*aHy, -aEyz*aHx (update_h.cl:149-154).
// mode 1: adds to Ebar (at the plane and its +d neighbour) and to Hbar for
// H_next; mode 2: adds the H_prev part.
template <class T>
__global__ __launch_bounds__(kBlock) void adj_obj(AdjArgs<T> q) {
    const Geom<T>& g = q.g;
    const int d = q.obj_axis;
    const int s0 = d == 0 ? g.sy : g.sx;
    const int s1 = d == 2 ? g.sy : g.sz;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= s0 * s1) return;
    const int c0 = t / s1, c1 = t - (t / s1) * s1;
    int x, y, z;
    if (d == 0) { x = q.obj_index; y = c0; z = c1; }
    else if (d == 1) { x = c0; y = q.obj_index; z = c1; }
    else { x = c0; y = c1; z = q.obj_index; }
    int xn = x, yn = y, zn = z;
    if (d == 0) xn = wrap(x + 1, g.sx);
    else if (d == 1) yn = wrap(y + 1, g.sy);
    else zn = wrap(z + 1, g.sz);
    const size_t i = ((size_t)x * g.sy + y) * g.sz + z;
    const size_t in = ((size_t)xn * g.sy + yn) * g.sz + zn;
    const int e1 = (d + 1) % 3, h1 = (d + 2) % 3;   // +aE_{e1 d} * aH_{h1}
    const int e2 = (d + 2) % 3, h2 = (d + 1) % 3;   // -aE_{e2 d} * aH_{h2}
    const T al = g.dt * (T)0.25;
    const T aE1 = q.En[e1][in] + q.En[e1][i];
    const T aE2 = q.En[e2][in] + q.En[e2][i];
    if (q.obj_mode == 1) {
        const T aH1 = q.Hn[h1][i] + q.Hp[h1][i];
        const T aH2 = q.Hn[h2][i] + q.Hp[h2][i];
        const T w1 = al * aH1, w2 = al * aH2;
        q.Eb[e1][i] += w1;
        q.Eb[e1][in] += w1;
        q.Eb[e2][i] -= w2;
        q.Eb[e2][in] -= w2;
        q.Hb[h1][i] += al * aE1;
        q.Hb[h2][i] -= al * aE2;
    } else {
        q.Hb[h1][i] += al * aE1;
        q.Hb[h2][i] -= al * aE2;
    }
}

// acc[0] += sum over the plane of one component, FP64, fixed order: one
// workgroup, each thread a fixed strided subset, then a tree in LDS.
template <class T>
__global__ __launch_bounds__(1024) void plane_sum(const T* f, int sx, int sy, int sz, int axis, int index,
                                                  double* acc) {
    __shared__ double part[1024];
    const int s0 = axis == 0 ? sy : sx;
    const int s1 = axis == 2 ? sy : sz;
    const int n = s0 * s1;
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) {
        const int c0 = t / s1, c1 = t - (t / s1) * s1;
        size_t i;
        if (axis == 0) i = ((size_t)index * sy + c0) * sz + c1;
        else if (axis == 1) i = ((size_t)c0 * sy + index) * sz + c1;
        else i = ((size_t)c0 * sy + c1) * sz + index;
        v += (double)f[i];
    }
    part[threadIdx.x] = v;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += part[0];
}

constexpr int kReduceBlocks = 1024;

// The same sum over many workgroups: workgroup b adds the plane cells
// [b * per, (b + 1) * per) (thread t: cells t, t + 256, ... of that range,
// in order), then its 256 sums pairwise, into partials[b]; reduce_accumulate
// then adds partials[0 .. nb) pairwise into acc[0].  A fixed order for a
// given plane size, so the result is reproducible.  A plane normal to z is
// one value per cache line; one workgroup reading all of them on one CU
// takes ~27 us on a 128 x 128 plane.
template <class T>
__global__ __launch_bounds__(256) void plane_partial(const T* f, int sx, int sy, int sz, int axis, int index, int per,
                                                     double* partials) {
    __shared__ double part[256];
    const int s0 = axis == 0 ? sy : sx;
    const int s1 = axis == 2 ? sy : sz;
    const int n = s0 * s1;
    const int lo = blockIdx.x * per;
    const int hi = min(lo + per, n);
    double v = 0.0;
    for (int t = lo + (int)threadIdx.x; t < hi; t += 256) {
        const int c0 = t / s1, c1 = t - (t / s1) * s1;
        size_t i;
        if (axis == 0) i = ((size_t)index * sy + c0) * sz + c1;
        else if (axis == 1) i = ((size_t)c0 * sy + index) * sz + c1;
        else i = ((size_t)c0 * sy + c1) * sz + index;
        v += (double)f[i];
    }
    part[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(1024) void reduce_accumulate(const double* partials, int nb, double* acc) {
    __shared__ double part[1024];
    part[threadIdx.x] = (int)threadIdx.x < nb ? partials[threadIdx.x] : 0.0;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += part[0];
}

// acc[0] += the plane sum through plane_partial (partials: kReduceBlocks
// doubles); declared before step_n_impl.
int plane_sum_many(const Plan& P, const void* field, int comp, int axis, int64_t index, double* partials,
                   double* acc, hipStream_t st) {
    const int64_t s[3] = {P.sx, P.sy, P.sz};
    const int64_t n = s[axis == 0 ? 1 : 0] * s[axis == 2 ? 1 : 2];
    // values per workgroup: a multiple of 256, at most kReduceBlocks workgroups
    int64_t per = (n + kReduceBlocks - 1) / kReduceBlocks;
    per = std::max<int64_t>(256, (per + 255) / 256 * 256);
    const int nb = (int)((n + per - 1) / per);
    if (P.d.dtype == FDTD_F32)
        hipLaunchKernelGGL(plane_partial<float>, dim3(nb), dim3(256), 0, st, (const float*)field + comp * P.N, P.sx,
                           P.sy, P.sz, axis, (int)index, (int)per, partials);
    else
        hipLaunchKernelGGL(plane_partial<double>, dim3(nb), dim3(256), 0, st, (const double*)field + comp * P.N,
                           P.sx, P.sy, P.sz, axis, (int)index, (int)per, partials);
    hipLaunchKernelGGL(reduce_accumulate, dim3(1), dim3(1024), 0, st, (const double*)partials, nb, acc);
    return check_launch("plane sum");
}

// Energy partials: each workgroup sums a fixed grid-stride subset of cells.
template <class T>
__global__ __launch_bounds__(256) void energy_partial(const T* E, const T* H, const T* Hp, const T* eps,
                                                      int64_t n3, double* partials) {
    __shared__ double part[256];
    double v = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n3; i += (int64_t)kReduceBlocks * 256) {
        const double e = (double)E[i];
        const double ep = (double)eps[i];
        if (isfinite(ep)) v += ep * e * e;
        v += (double)H[i] * (double)Hp[i];
    }
    part[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(1024) void reduce_final(const double* partials, int n, double* out) {
    __shared__ double part[1024];
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 1024) v += partials[t];
    part[threadIdx.x] = v;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

template <class T>
int launch_adjoint(const Plan& P, const fdtd_adjoint_bufs* b, hipStream_t st) {
    AdjArgs<T> q;
    memset(&q, 0, sizeof(q));
    fdtd_fields f;
    memset(&f, 0, sizeof(f));
    f.coef = (void*)b->coef;
    fill_geom<T>(P, P.le, &f, q.g);
    const int64_t N = P.N;
    comp3c<T>(b->E_prev, N, q.Ep);
    comp3c<T>(b->E_next, N, q.En);
    comp3c<T>(b->H_prev, N, q.Hp);
    comp3c<T>(b->H_next, N, q.Hn);
    comp3c<T>(b->eps, N, q.eps);
    comp3<T>(b->Eb, N, q.Eb);
    comp3<T>(b->Hb, N, q.Hb);
    for (int k = 0; k < 3; ++k) q.grad[k] = b->eps_grad ? b->eps_grad + k * N : nullptr;
    const bool whole = b->grad_hi[0] == 0 && b->grad_hi[1] == 0 && b->grad_hi[2] == 0;
    const int64_t s[3] = {P.sx, P.sy, P.sz};
    for (int k = 0; k < 3; ++k) {
        q.glo[k] = whole ? 0 : b->grad_lo[k];
        q.ghi[k] = whole ? (int)s[k] : b->grad_hi[k];
    }
    q.src_comp = -1;
    if (b->src_index >= 0) {
        const int64_t comp = b->src_index / N;
        const int64_t i = b->src_index - comp * N;
        q.src_comp = (int)comp;
        q.src_x = (int)(i / P.plane);
        q.src_p = (int)(i - (int64_t)q.src_x * P.plane);
        q.src_amp = (T)b->src_amp;
    }
    q.obj_axis = b->obj_axis;
    q.obj_index = (int)b->obj_index;
    const dim3 grid((P.plane + kBlock - 1) / kBlock, P.sx);
    int obj_cells = 0;
    if (b->obj_axis >= 0) {
        const int64_t s[3] = {P.sx, P.sy, P.sz};
        obj_cells = (int)(P.N / s[b->obj_axis]);
    }
    const dim3 ogrid((obj_cells + kBlock - 1) / kBlock);
    const int parts = (b->parts == 0) ? 3 : b->parts;
    if (parts & 1) {
    // 1. objective: Ebar and the H_next part of Hbar
    if (obj_cells) {
        q.obj_mode = 1;
        hipLaunchKernelGGL(adj_obj<T>, ogrid, dim3(kBlock), 0, st, q);
    }
    // 2. H step reversed
    q.pin = (const T*)b->psib_h;
    q.pout = (T*)b->psib_h_out;
    hipLaunchKernelGGL(adj_h<T>, grid, dim3(kBlock), 0, st, q);
    // 3. objective: the H_prev part of Hbar
    if (obj_cells) {
        q.obj_mode = 2;
        hipLaunchKernelGGL(adj_obj<T>, ogrid, dim3(kBlock), 0, st, q);
    }
    }
    if (parts & 2) {
        // 4. E step reversed (the source adds a constant: no adjoint term)
        q.pin = (const T*)b->psib_e;
        q.pout = (T*)b->psib_e_out;
        hipLaunchKernelGGL(adj_e<T>, grid, dim3(kBlock), 0, st, q);
    }
    return check_launch("fdtd_adjoint_step");
}

}  // namespace

extern "C" {

int fdtd_adjoint_step(const void* plan, const fdtd_adjoint_bufs* b, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (!b || !b->E_prev || !b->E_next || !b->H_prev || !b->H_next || !b->eps || !b->Eb || !b->Hb)
        return fail(2, "adjoint: E_prev, E_next, H_prev, H_next, eps, Eb and Hb are required");
    if (b->parts < 0 || b->parts > 3) return fail(3, "adjoint: parts %d not in 0..3", b->parts);
    if (P.d.nonuniform) return fail(4, "adjoint: non-uniform grids are not supported");
    if (P.d.bloch_mask) return fail(4, "adjoint: Bloch boundaries are not supported");
    if (P.d.n_face > 0) {
        if (!b->coef || !b->psib_e || !b->psib_e_out || !b->psib_h || !b->psib_h_out)
            return fail(2, "adjoint: coef and the four adjoint psi blocks are required with CPML");
        if (b->psib_e == b->psib_e_out || b->psib_h == b->psib_h_out)
            return fail(2, "adjoint: the adjoint psi blocks must be ping-ponged (in != out)");
        for (int k = 0; k < P.d.n_face; ++k)
            for (int l = k + 1; l < P.d.n_face; ++l)
                if (P.d.face[k].axis == P.d.face[l].axis && P.face[k].hi > P.face[l].lo &&
                    P.face[l].hi > P.face[k].lo)
                    return fail(4, "adjoint: overlapping CPML slabs on axis %d", P.d.face[k].axis);
    }
    if (b->obj_axis > 2 || b->obj_axis < -1) return fail(3, "adjoint: obj_axis %d not in -1..2", b->obj_axis);
    if (b->obj_axis >= 0) {
        const int64_t s[3] = {P.sx, P.sy, P.sz};
        if (b->obj_index < 0 || b->obj_index >= s[b->obj_axis])
            return fail(3, "adjoint: obj_index %d out of range", (int)b->obj_index);
    }
    if (b->src_index >= 3 * P.N) return fail(3, "adjoint: src_index out of range");
    if (b->grad_hi[0] != 0 || b->grad_hi[1] != 0 || b->grad_hi[2] != 0) {
        const int64_t s[3] = {P.sx, P.sy, P.sz};
        for (int k = 0; k < 3; ++k)
            if (b->grad_lo[k] < 0 || b->grad_lo[k] >= b->grad_hi[k] || b->grad_hi[k] > s[k])
                return fail(3, "adjoint: gradient box [%d, %d) on axis %d not within [0, %lld)", b->grad_lo[k],
                            b->grad_hi[k], k, (long long)s[k]);
    }
    hipStream_t st = (hipStream_t)stream;
    if (P.d.dtype == FDTD_F32) return launch_adjoint<float>(P, b, st);
    return launch_adjoint<double>(P, b, st);
}

int fdtd_plane_sum(const void* plan, const void* field, int32_t comp, int32_t axis, int64_t index, double* acc,
                   void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (!field || !acc) return fail(2, "field and acc are required");
    if (comp < 0 || comp > 2 || axis < 0 || axis > 2) return fail(3, "comp/axis out of range");
    const int64_t s[3] = {P.sx, P.sy, P.sz};
    if (index < 0 || index >= s[axis]) return fail(3, "plane index %lld out of range", (long long)index);
    hipStream_t st = (hipStream_t)stream;
    if (P.d.dtype == FDTD_F32)
        hipLaunchKernelGGL(plane_sum<float>, dim3(1), dim3(1024), 0, st, (const float*)field + comp * P.N, P.sx,
                           P.sy, P.sz, axis, (int)index, acc);
    else
        hipLaunchKernelGGL(plane_sum<double>, dim3(1), dim3(1024), 0, st, (const double*)field + comp * P.N,
                           P.sx, P.sy, P.sz, axis, (int)index, acc);
    return check_launch("fdtd_plane_sum");
}

int64_t fdtd_reduce_partials(const void* plan) { return plan ? kReduceBlocks : -1; }

}  // extern "C"

// Scratch of fdtd_energy_noalloc: the FP64 partial sums of energy_partial,
// one per workgroup (8 KiB).  A file-scope device array, so the entry point
// needs no allocation; each GPU the library runs on holds its own copy until
// the process ends.  Host code must not pass the array's name to a kernel
// (on the host it is only a stand-in for the symbol): it uses the device
// address hipGetSymbolAddress returns, kept per device.
constexpr int kScratchPartials = 1024;
static __device__ __align__(256) double scratch[kScratchPartials];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};

// The current device's address of `scratch`, looked up once per device with
// that device current.
int scratch_address(double** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return fail(4, "hipGetDevice failed");
    if (dev < 0 || dev >= kScratchDevices) return fail(4, "device %d beyond the %d supported", dev, kScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(4, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

int fdtd_energy(const void* plan, const void* E, const void* H, const void* H_prev, const void* eps,
                double* partials, double* out, void* stream) {
    if (!plan) return fail(1, "plan is NULL");
    const Plan& P = *(const Plan*)plan;
    if (!E || !H || !H_prev || !eps || !partials || !out) return fail(2, "all buffers are required");
    hipStream_t st = (hipStream_t)stream;
    if (P.d.dtype == FDTD_F32)
        hipLaunchKernelGGL(energy_partial<float>, dim3(kReduceBlocks), dim3(256), 0, st, (const float*)E,
                           (const float*)H, (const float*)H_prev, (const float*)eps, 3 * P.N, partials);
    else
        hipLaunchKernelGGL(energy_partial<double>, dim3(kReduceBlocks), dim3(256), 0, st, (const double*)E,
                           (const double*)H, (const double*)H_prev, (const double*)eps, 3 * P.N, partials);
    hipLaunchKernelGGL(reduce_final, dim3(1), dim3(1024), 0, st, partials, kReduceBlocks, out);
    return check_launch("fdtd_energy");
}

int fdtd_energy_noalloc(const void* plan, const void* E, const void* H, const void* H_prev, const void* eps,
                        double* out, void* stream) {
    static_assert(kReduceBlocks <= kScratchPartials, "the energy reduction must fit the scratch array");
    if (!plan) return fail(1, "plan is NULL");
    if (fdtd_reduce_partials(plan) > kScratchPartials)
        return fail(3, "the reduction needs %lld partials, the scratch holds %d",
                    (long long)fdtd_reduce_partials(plan), kScratchPartials);
    double* partials = nullptr;
    if (int rc = scratch_address(&partials)) return rc;
    // energy_partial writes every partial before reduce_final reads it, so
    // nothing needs clearing; calls that overlap on one GPU share the array
    // and race
    return fdtd_energy(plan, E, H, H_prev, eps, partials, out, stream);
}

}  // extern "C"
