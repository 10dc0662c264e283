// CPU oracle for the FDTD time-stepping hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// library built from it.  The product path (opencl_fdtd_amd) never calls it.
//
// It restates, cell by cell and in the same floating-point operation order,
// the per-work-item bodies of the reference's Jinja2/OpenCL templates:
//   - index decode + periodic wrap        opencl_fdtd/kernels/common.cl:16-85
//   - E update (+ CPML psi_E)             opencl_fdtd/kernels/update_e.cl:20-107
//   - H update (+ CPML psi_H, oS)         opencl_fdtd/kernels/update_h.cl:22-154
//   - S from oS                           opencl_fdtd/kernels/update_s.cl:21-36
//   - field source J                      opencl_fdtd/kernels/update_j.cl:14-32
// The reference dataflow is kept: three passes per step, with the 6N oS
// intermediate written by the H pass and read by the S pass.
//
// Bloch seams (the reference's Bloch block never renders, so its intended
// semantics is defined here): the neighbour read across the seam of a Bloch
// axis r is  a_r * own[j] + b_r * partner[j]  where j is the periodic wrap
// index and (a_r, b_r) are supplied per call.  With a_r = 1, b_r = 0 this is
// exactly the reference's plain periodic wrap.
//
// Compile with -ffp-contract=off: every product and sum is rounded
// separately, in the type of the field (float or double).

#include <cstdint>
#include <cstddef>
#include <cmath>
#include <omp.h>

extern "C" {

#define ORACLE_MAX_FACES 12

struct oracle_face {
    int32_t axis;       // 0 = x, 1 = y, 2 = z
    int32_t polarity;   // 0 = 'n' (low side), 1 = 'p' (high side)
    int32_t thickness;
    int32_t pad_;
    const void* p0e; const void* p1e; const void* p2e;   // length = thickness
    const void* p0h; const void* p1h; const void* p2h;
    void* psi_eu; void* psi_ev;   // psi_E for the two transverse axes (u < v)
    void* psi_hu; void* psi_hv;   // psi_H
};

struct oracle_grid {
    int64_t sx, sy, sz;
    int32_t dtype;          // 0 = float32, 1 = float64
    int32_t n_face;
    oracle_face face[ORACLE_MAX_FACES];
    double inv_d_uniform;   // (double)1 / dx, used when inv_de/inv_dh are null
    const void* inv_de[3];  // per-index 1/dx at E locations (H-step), or null
    const void* inv_dh[3];  // per-index 1/dx at H locations (E-step), or null
    int32_t nthreads;       // OpenMP threads, <= 0: runtime default
    int32_t pad2_;
};

}  // extern "C"

namespace {

inline int nthr(const oracle_grid* g) {
    return g->nthreads > 0 ? g->nthreads : omp_get_max_threads();
}

struct Geo {
    int64_t s[3];
    int64_t N;
    int64_t st[3];   // strides: x = sy*sz, y = sz, z = 1
};

inline Geo make_geo(const oracle_grid* g) {
    Geo q;
    q.s[0] = g->sx; q.s[1] = g->sy; q.s[2] = g->sz;
    q.N = g->sx * g->sy * g->sz;
    q.st[0] = g->sy * g->sz; q.st[1] = g->sz; q.st[2] = 1;
    return q;
}

// Periodic offsets, common.cl:75-85.
inline int64_t off_minus(const Geo& q, int r, int64_t c) {
    return (c == 0) ? (q.s[r] - 1) * q.st[r] : -q.st[r];
}
inline int64_t off_plus(const Geo& q, int r, int64_t c) {
    return (c == q.s[r] - 1) ? -(q.s[r] - 1) * q.st[r] : q.st[r];
}

// Slab membership and parameter index, update_e.cl:66-76 / update_h.cl:97-107.
// The reference compares size_t values; s - T wraps when T > s, making the
// 'p' test false.
inline bool in_slab(const oracle_face& f, const Geo& q, const int64_t* c, int64_t* ir) {
    const int r = f.axis;
    const int64_t T = f.thickness;
    if (f.polarity == 0) {
        if (c[r] < T) { *ir = c[r]; return true; }
        return false;
    }
    if (T > q.s[r]) return false;
    if (q.s[r] > c[r] && c[r] >= q.s[r] - T) { *ir = (q.s[r] - 1) - c[r]; return true; }
    return false;
}

inline void transverse(int r, int* u, int* v) {
    if (r == 0) { *u = 1; *v = 2; }
    else if (r == 1) { *u = 0; *v = 2; }
    else { *u = 0; *v = 1; }
}

template <class T>
void update_e(const oracle_grid* g, T* E, const T* H, const T* P, const T* eps, double dt_d,
              int seam_mask, const double* seam_ab) {
    const Geo q = make_geo(g);
    const int64_t N = q.N;
    const T dt = (T)dt_d;
    const T inv_u = (T)g->inv_d_uniform;
    T sa[3], sb[3];
    for (int r = 0; r < 3; r++) {
        sa[r] = seam_ab ? (T)seam_ab[r] : (T)1;
        sb[r] = seam_ab ? (T)seam_ab[3 + r] : (T)0;
    }
    const T* inv_dh[3] = {(const T*)g->inv_dh[0], (const T*)g->inv_dh[1], (const T*)g->inv_dh[2]};
#pragma omp parallel for schedule(static) num_threads(nthr(g))
    for (int64_t x = 0; x < q.s[0]; x++) {
        for (int64_t y = 0; y < q.s[1]; y++) {
            for (int64_t z = 0; z < q.s[2]; z++) {
                const int64_t c[3] = {x, y, z};
                const int64_t i = z + q.s[2] * (y + q.s[1] * x);
                int64_t m[3];
                for (int r = 0; r < 3; r++) m[r] = off_minus(q, r, c[r]);
                // H_a at i - e_b; across a Bloch seam mixed with the partner field.
                auto hb = [&](int a, int b) -> T {
                    T val = H[a * N + i + m[b]];
                    if (((seam_mask >> b) & 1) && c[b] == 0)
                        val = sa[b] * val + sb[b] * P[a * N + i + m[b]];
                    return val;
                };
                T inv[3];
                for (int r = 0; r < 3; r++) inv[r] = inv_dh[r] ? inv_dh[r][c[r]] : inv_u;
                // dH[a][b] = d(H_a)/d(b), backward difference (update_e.cl:39-46)
                T dH[3][3] = {};
                dH[0][1] = (H[0 * N + i] - hb(0, 1)) * inv[1];
                dH[0][2] = (H[0 * N + i] - hb(0, 2)) * inv[2];
                dH[1][0] = (H[1 * N + i] - hb(1, 0)) * inv[0];
                dH[1][2] = (H[1 * N + i] - hb(1, 2)) * inv[2];
                dH[2][0] = (H[2 * N + i] - hb(2, 0)) * inv[0];
                dH[2][1] = (H[2 * N + i] - hb(2, 1)) * inv[1];
                T pE[3] = {(T)0, (T)0, (T)0};
                for (int k = 0; k < g->n_face; k++) {
                    const oracle_face& f = g->face[k];
                    int64_t ir;
                    if (!in_slab(f, q, c, &ir)) continue;
                    const int r = f.axis;
                    int u, v;
                    transverse(r, &u, &v);
                    const int64_t ip = c[v] + c[u] * q.s[v] + ir * q.s[v] * q.s[u];   // update_e.cl:92
                    const T p0 = ((const T*)f.p0e)[ir];
                    const T p1 = ((const T*)f.p1e)[ir];
                    const T p2 = ((const T*)f.p2e)[ir];
                    T* psu = (T*)f.psi_eu;
                    T* psv = (T*)f.psi_ev;
                    dH[v][r] *= p2;
                    dH[u][r] *= p2;
                    psu[ip] = p0 * psu[ip] + p1 * dH[v][r];
                    psv[ip] = p0 * psv[ip] + p1 * dH[u][r];
                    if (r != 1) { pE[u] -= psu[ip]; pE[v] += psv[ip]; }
                    else        { pE[u] += psu[ip]; pE[v] -= psv[ip]; }
                }
                // update_e.cl:105-107
                E[0 * N + i] += dt / eps[0 * N + i] * (dH[2][1] - dH[1][2] + pE[0]);
                E[1 * N + i] += dt / eps[1 * N + i] * (dH[0][2] - dH[2][0] + pE[1]);
                E[2 * N + i] += dt / eps[2 * N + i] * (dH[1][0] - dH[0][1] + pE[2]);
            }
        }
    }
}

template <class T>
void update_h(const oracle_grid* g, const T* E, T* H, const T* P, T* oS, double dt_d,
              int seam_mask, const double* seam_ab) {
    const Geo q = make_geo(g);
    const int64_t N = q.N;
    const T dt = (T)dt_d;
    const T inv_u = (T)g->inv_d_uniform;
    T sa[3], sb[3];
    for (int r = 0; r < 3; r++) {
        sa[r] = seam_ab ? (T)seam_ab[r] : (T)1;
        sb[r] = seam_ab ? (T)seam_ab[3 + r] : (T)0;
    }
    const T* inv_de[3] = {(const T*)g->inv_de[0], (const T*)g->inv_de[1], (const T*)g->inv_de[2]};
#pragma omp parallel for schedule(static) num_threads(nthr(g))
    for (int64_t x = 0; x < q.s[0]; x++) {
        for (int64_t y = 0; y < q.s[1]; y++) {
            for (int64_t z = 0; z < q.s[2]; z++) {
                const int64_t c[3] = {x, y, z};
                const int64_t i = z + q.s[2] * (y + q.s[1] * x);
                int64_t pp[3];
                for (int r = 0; r < 3; r++) pp[r] = off_plus(q, r, c[r]);
                // E_a at i + e_b; across a Bloch seam mixed with the partner field.
                auto ef = [&](int a, int b) -> T {
                    T val = E[a * N + i + pp[b]];
                    if (((seam_mask >> b) & 1) && c[b] == q.s[b] - 1)
                        val = sa[b] * val + sb[b] * P[a * N + i + pp[b]];
                    return val;
                };
                T inv[3];
                for (int r = 0; r < 3; r++) inv[r] = inv_de[r] ? inv_de[r][c[r]] : inv_u;
                const T ex = E[0 * N + i], ey = E[1 * N + i], ez = E[2 * N + i];
                const T ex_y = ef(0, 1), ex_z = ef(0, 2);
                const T ey_x = ef(1, 0), ey_z = ef(1, 2);
                const T ez_x = ef(2, 0), ez_y = ef(2, 1);
                // dE[a][b] = d(E_a)/d(b), forward difference (update_h.cl:36-43)
                T dE[3][3] = {};
                dE[0][1] = (ex_y - ex) * inv[1];
                dE[0][2] = (ex_z - ex) * inv[2];
                dE[1][0] = (ey_x - ey) * inv[0];
                dE[1][2] = (ey_z - ey) * inv[2];
                dE[2][0] = (ez_x - ez) * inv[0];
                dE[2][1] = (ez_y - ez) * inv[1];
                T pH[3] = {(T)0, (T)0, (T)0};
                for (int k = 0; k < g->n_face; k++) {
                    const oracle_face& f = g->face[k];
                    int64_t ir;
                    if (!in_slab(f, q, c, &ir)) continue;
                    const int r = f.axis;
                    int u, v;
                    transverse(r, &u, &v);
                    const int64_t ip = c[v] + c[u] * q.s[v] + ir * q.s[v] * q.s[u];   // update_h.cl:108
                    const T p0 = ((const T*)f.p0h)[ir];
                    const T p1 = ((const T*)f.p1h)[ir];
                    const T p2 = ((const T*)f.p2h)[ir];
                    T* psu = (T*)f.psi_hu;
                    T* psv = (T*)f.psi_hv;
                    dE[v][r] *= p2;
                    dE[u][r] *= p2;
                    psu[ip] = p0 * psu[ip] + p1 * dE[v][r];
                    psv[ip] = p0 * psv[ip] + p1 * dE[u][r];
                    if (r != 1) { pH[u] += psu[ip]; pH[v] -= psv[ip]; }
                    else        { pH[u] -= psu[ip]; pH[v] += psv[ip]; }
                }
                const T hx0 = H[0 * N + i], hy0 = H[1 * N + i], hz0 = H[2 * N + i];
                // update_h.cl:129-131
                const T hx1 = hx0 - dt * (dE[2][1] - dE[1][2] - pH[0]);
                const T hy1 = hy0 - dt * (dE[0][2] - dE[2][0] - pH[1]);
                const T hz1 = hz0 - dt * (dE[1][0] - dE[0][1] - pH[2]);
                H[0 * N + i] = hx1;
                H[1 * N + i] = hy1;
                H[2 * N + i] = hz1;
                if (oS) {
                    // update_h.cl:63-70, 135-154
                    const T aExy = ex_y + ex, aExz = ex_z + ex;
                    const T aEyx = ey_x + ey, aEyz = ey_z + ey;
                    const T aEzx = ez_x + ez, aEzy = ez_y + ez;
                    const T aHx = hx1 + hx0, aHy = hy1 + hy0, aHz = hz1 + hz0;
                    oS[0 * N + i] = aEyx * aHz;      // oSxy
                    oS[3 * N + i] = -aEzx * aHy;     // oSxz
                    oS[1 * N + i] = aEzy * aHx;      // oSyz
                    oS[4 * N + i] = -aExy * aHz;     // oSyx
                    oS[2 * N + i] = aExz * aHy;      // oSzx
                    oS[5 * N + i] = -aEyz * aHx;     // oSzy
                }
            }
        }
    }
}

template <class T>
void update_s(const oracle_grid* g, const T* oS, T* S, double dt_d) {
    const Geo q = make_geo(g);
    const int64_t N = q.N;
    const T dt = (T)dt_d;
    const T s_factor = dt * (T)0.125;   // update_s.cl:33
    const T* oSxy = oS + 0 * N; const T* oSyz = oS + 1 * N; const T* oSzx = oS + 2 * N;
    const T* oSxz = oS + 3 * N; const T* oSyx = oS + 4 * N; const T* oSzy = oS + 5 * N;
#pragma omp parallel for schedule(static) num_threads(nthr(g))
    for (int64_t x = 0; x < q.s[0]; x++) {
        for (int64_t y = 0; y < q.s[1]; y++) {
            for (int64_t z = 0; z < q.s[2]; z++) {
                const int64_t i = z + q.s[2] * (y + q.s[1] * x);
                const int64_t mx = off_minus(q, 0, x), my = off_minus(q, 1, y), mz = off_minus(q, 2, z);
                // update_s.cl:34-36, summed left to right
                S[0 * N + i] = (oSxy[i] + oSxz[i] + oSxy[i + my] + oSxz[i + mz]) * s_factor;
                S[1 * N + i] = (oSyz[i] + oSyx[i] + oSyz[i + mz] + oSyx[i + mx]) * s_factor;
                S[2 * N + i] = (oSzx[i] + oSzy[i] + oSzx[i + mx] + oSzy[i + my]) * s_factor;
            }
        }
    }
}

template <class T>
void update_j(const oracle_grid* g, T* E, const T* Jr, const T* Ji, double c_d, double s_d,
              const int64_t* box) {
    const Geo q = make_geo(g);
    const int64_t N = q.N;
    const T cc = (T)c_d, ss = (T)s_d;
    // update_j.cl:21-32: cells with min <= coord < max on every axis
#pragma omp parallel for schedule(static) num_threads(nthr(g))
    for (int64_t x = 0; x < q.s[0]; x++) {
        for (int64_t y = 0; y < q.s[1]; y++) {
            for (int64_t z = 0; z < q.s[2]; z++) {
                if (x < box[0] || y < box[2] || z < box[4]) continue;
                if (x >= box[1] || y >= box[3] || z >= box[5]) continue;
                const int64_t i = z + q.s[2] * (y + q.s[1] * x);
                for (int a = 0; a < 3; a++)
                    E[a * N + i] += cc * Jr[a * N + i] + ss * Ji[a * N + i];
            }
        }
    }
}

}  // namespace

extern "C" {

int oracle_abi_version(void) { return 1; }

int oracle_max_faces(void) { return ORACLE_MAX_FACES; }

// seam_ab: 6 doubles (a_x, a_y, a_z, b_x, b_y, b_z) or null; seam_mask bit r
// marks Bloch axis r.
int oracle_update_e(const oracle_grid* g, void* E, const void* H, const void* P, const void* eps,
                    double dt, int seam_mask, const double* seam_ab) {
    if (!g || !E || !H || !eps || g->n_face > ORACLE_MAX_FACES) return 1;
    if (seam_mask && !P) return 2;
    if (g->dtype == 0)
        update_e<float>(g, (float*)E, (const float*)H, (const float*)P, (const float*)eps, dt, seam_mask, seam_ab);
    else
        update_e<double>(g, (double*)E, (const double*)H, (const double*)P, (const double*)eps, dt, seam_mask, seam_ab);
    return 0;
}

int oracle_update_h(const oracle_grid* g, const void* E, void* H, const void* P, void* oS,
                    double dt, int seam_mask, const double* seam_ab) {
    if (!g || !E || !H || g->n_face > ORACLE_MAX_FACES) return 1;
    if (seam_mask && !P) return 2;
    if (g->dtype == 0)
        update_h<float>(g, (const float*)E, (float*)H, (const float*)P, (float*)oS, dt, seam_mask, seam_ab);
    else
        update_h<double>(g, (const double*)E, (double*)H, (const double*)P, (double*)oS, dt, seam_mask, seam_ab);
    return 0;
}

int oracle_update_s(const oracle_grid* g, const void* oS, void* S, double dt) {
    if (!g || !oS || !S) return 1;
    if (g->dtype == 0) update_s<float>(g, (const float*)oS, (float*)S, dt);
    else update_s<double>(g, (const double*)oS, (double*)S, dt);
    return 0;
}

int oracle_update_j(const oracle_grid* g, void* E, const void* Jr, const void* Ji, double c, double s,
                    const int64_t* box) {
    if (!g || !E || !Jr || !Ji || !box) return 1;
    if (g->dtype == 0) update_j<float>(g, (float*)E, (const float*)Jr, (const float*)Ji, c, s, box);
    else update_j<double>(g, (double*)E, (const double*)Jr, (const double*)Ji, c, s, box);
    return 0;
}

}  // extern "C"
