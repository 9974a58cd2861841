// Device-side initialisation: the material model (sord/material_model.f90:
// 18-169), the node masses and viscosity of the resampling
// (sord/material_resample.f90:6-55) and the volumes of rectilinear cells
// (sord/grid_generation.f90:183-198), computed from the input arrays on the
// device.  Every kernel performs the operations of the host restatement
// (sord_init.inc) in the same order and the same precision, with no
// contraction of products and sums into fused multiply-adds, so the arrays
// agree with the host's bit for bit.
#pragma clang fp contract(off)

#include "sord_dev.h"
#include "sord_launch.h"

namespace sord {

namespace {

// element q of the mesh (one thread each) as 1-based (j, k, l)
__device__ __forceinline__ bool cell_of(long q, int &j, int &k, int &l, const int3 nm) {
    if (q >= (long)nm.x * nm.y * nm.z) return false;
    j = 1 + (int)(q % nm.x);
    k = 1 + (int)((q / nm.x) % nm.y);
    l = 1 + (int)(q / ((long)nm.x * nm.y));
    return true;
}

__device__ __forceinline__ bool inside(int j, int k, int l, int3 a, int3 b) {
    return j >= a.x && j <= b.x && k >= a.y && k <= b.y && l >= a.z && l <= b.z;
}

// min(x, hi) then max(., lo) as the host's std::min / std::max evaluate them
template <class T>
__device__ __forceinline__ T limit(T x, T lo, T hi) {
    if (lo > 0 && hi > 0) {
        x = hi < x ? hi : x;
        return x < lo ? lo : x;
    }
    if (lo > 0) return x < lo ? lo : x;
    if (hi > 0) return hi < x ? hi : x;
    return x;
}

}  // namespace

// cell volumes of a rectilinear mesh: the product of the three edge lengths
// (sign flipped with the sign of dx dy dz), zero outside the cells and on the
// fault's zero-width cell
template <class T>
__global__ void k_init_vc(T *vc, Grid<T> g, const T *hx, const T *hy, const T *hz, int3 c1, int3 c2, int flip,
                          int fa, int irup) {
    const int3 nm = make_int3(g.nm[0], g.nm[1], g.nm[2]);
    int j, k, l;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (!cell_of(q, j, k, l, nm)) return;
    T v = T(0);
    if (inside(j, k, l, c1, c2)) {
        v = hx[j] * hy[k] * hz[l];
        if (flip) v = -v;
    }
    const int c3[3] = {j, k, l};
    if (fa >= 0 && c3[fa] == irup) v = T(0);
    vc[g.at(j, k, l)] = v;
}

// the cell material: limits, viscosity from vdamp, zero outside the cells
// (halo), then mu = rho vs^2, lambda = rho vp^2 - 2 mu and Poisson's ratio.
// In: rho, vp, vs, gam as read; out: rho, vp, vs after the limits (rc, s1,
// s2), gam, lam, mu, nu.  vp, vs may share storage with lam, mu.
template <class T>
__global__ void k_init_cells(Grid<T> g, MatInit<T> m, const T *rho, const T *vp, const T *vs, T *gam, T *rc, T *s1,
                             T *s2, T *lam, T *mu, T *nu) {
    const int3 nm = make_int3(g.nm[0], g.nm[1], g.nm[2]);
    int j, k, l;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (!cell_of(q, j, k, l, nm)) return;
    const long o = g.at(j, k, l);
    T r = limit(rho[o], m.rho_min, m.rho_max);
    T a = limit(vp[o], m.vp_min, m.vp_max);
    T b = limit(vs[o], m.vs_min, m.vs_max);
    T y = gam[o];
    if (m.vdamp > 0) y = b == 0 ? T(0) : m.vdamp / b;
    y = limit(y, m.gam_min, m.gam_max);
    // the halo of set_halo (sord/utilities.f90:103-118): singleton axes keep theirs
    if ((nm.x > 1 && (j < m.h1.x || j > m.h2.x)) || (nm.y > 1 && (k < m.h1.y || k > m.h2.y)) ||
        (nm.z > 1 && (l < m.h1.z || l > m.h2.z)))
        r = a = b = y = T(0);
    const T u = r * b * b;
    const T la = r * a * a - T(2) * u;
    const T s = la + u;
    rc[o] = r, s1[o] = a, s2[o] = b, gam[o] = y;
    lam[o] = la, mu[o] = u;
    nu[o] = s == 0 ? T(0) : T(0.5) * la / s;
}

// minima and maxima of seven arrays over a box, one row of 14 per block
// (minima, then maxima); the host reduces the rows
template <class T>
__global__ void __launch_bounds__(256) k_init_minmax(Grid<T> g, MatArrays<T> a, int3 lo, int3 hi, T *part) {
    __shared__ T sh[14][256];
    T mn[7], mx[7];
#pragma unroll
    for (int i = 0; i < 7; i++) mn[i] = MatInit<T>::big(), mx[i] = -MatInit<T>::big();
    const int3 n = make_int3(hi.x - lo.x + 1, hi.y - lo.y + 1, hi.z - lo.z + 1);
    const long tot = n.x > 0 && n.y > 0 && n.z > 0 ? (long)n.x * n.y * n.z : 0;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < tot; q += (long)gridDim.x * blockDim.x) {
        const int j = lo.x + (int)(q % n.x), k = lo.y + (int)((q / n.x) % n.y), l = lo.z + (int)(q / ((long)n.x * n.y));
        const long o = g.at(j, k, l);
#pragma unroll
        for (int i = 0; i < 7; i++) {
            const T v = a.p[i][o];
            mn[i] = v < mn[i] ? v : mn[i];
            mx[i] = mx[i] < v ? v : mx[i];
        }
    }
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < 7; i++) sh[i][t] = mn[i], sh[7 + i][t] = mx[i];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int i = 0; i < 7; i++) {
                const T b0 = sh[i][t + w], b1 = sh[7 + i][t + w];
                sh[i][t] = b0 < sh[i][t] ? b0 : sh[i][t];
                sh[7 + i][t] = sh[7 + i][t] < b1 ? b1 : sh[7 + i][t];
            }
        __syncthreads();
    }
    if (t < 14) part[blockIdx.x * 14 + t] = sh[t][0];
}

// hourglass stiffness (sord/material_model.f90:130): mu (lam + mu) s / (lam + 2 mu)
template <class T>
__global__ void k_init_yy(Grid<T> g, T *yy, const T *lam, const T *mu, T s) {
    const int3 nm = make_int3(g.nm[0], g.nm[1], g.nm[2]);
    int j, k, l;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (!cell_of(q, j, k, l, nm)) return;
    const long o = g.at(j, k, l);
    const T la = lam[o], u = mu[o];
    const T r = la + T(2) * u;
    yy[o] = r == 0 ? T(0) : u * (la + u) * s / r;
}

// out = a b (b given) or a c, every element of the mesh
template <class T>
__global__ void k_init_scale(Grid<T> g, T *out, const T *a, const T *b, T c) {
    const int3 nm = make_int3(g.nm[0], g.nm[1], g.nm[2]);
    int j, k, l;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (!cell_of(q, j, k, l, nm)) return;
    const long o = g.at(j, k, l);
    out[o] = b ? a[o] * b[o] : a[o] * c;
}

// the average of the 8 cells around each node of [n1, n2], zero elsewhere;
// with inv, its inverse where it is not zero
template <class T>
__global__ void k_init_nodeavg(Grid<T> g, T *out, const T *in, int3 n1, int3 n2, int inv) {
    const int3 nm = make_int3(g.nm[0], g.nm[1], g.nm[2]);
    int j, k, l;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (!cell_of(q, j, k, l, nm)) return;
    T v = T(0);
    if (inside(j, k, l, n1, n2)) {
        auto f = [&](int a, int b, int c) { return in[g.at(a, b, c)]; };
        v = T(0.125) * (f(j, k, l) + f(j - 1, k - 1, l - 1) + f(j, k - 1, l - 1) + f(j - 1, k, l) +
                        f(j - 1, k, l - 1) + f(j, k - 1, l) + f(j - 1, k - 1, l) + f(j, k, l - 1));
    }
    if (inv && v != 0) v = T(1) / v;
    out[g.at(j, k, l)] = v;
}

// a = 1 / a where a is not zero
template <class T>
__global__ void k_init_recip(Grid<T> g, T *a) {
    const int3 nm = make_int3(g.nm[0], g.nm[1], g.nm[2]);
    int j, k, l;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (!cell_of(q, j, k, l, nm)) return;
    const long o = g.at(j, k, l);
    const T v = a[o];
    if (v != 0) a[o] = T(1) / v;
}

// f = r outside [i1, i2] (set_halo, sord/utilities.f90:103-118; singleton axes
// keep their values), and on the plane idx of axis pa if pa >= 0
template <class T>
__global__ void k_init_halo(Grid<T> g, T *f, T r, int3 i1, int3 i2, int pa, int idx) {
    const int3 nm = make_int3(g.nm[0], g.nm[1], g.nm[2]);
    int j, k, l;
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (!cell_of(q, j, k, l, nm)) return;
    const int i3x = min(i1.x, nm.x + 1), i3y = min(i1.y, nm.y + 1), i3z = min(i1.z, nm.z + 1);
    const int i4x = max(i2.x, 0), i4y = max(i2.y, 0), i4z = max(i2.z, 0);
    bool h = (nm.x > 1 && (j < i3x || j > i4x)) || (nm.y > 1 && (k < i3y || k > i4y)) ||
             (nm.z > 1 && (l < i3z || l > i4z));
    const int c3[3] = {j, k, l};
    if (pa >= 0 && c3[pa] == idx) h = true;
    if (h) f[g.at(j, k, l)] = r;
}

namespace {
inline unsigned blocks_of(const int nm[3]) {
    const long n = (long)nm[0] * nm[1] * nm[2];
    return (unsigned)((n + 255) / 256);
}
inline int3 i3(const int a[3]) { return make_int3(a[0], a[1], a[2]); }
}  // namespace

template <class T>
void launch_init_vc(T *vc, const Grid<T> &g, const T *hx, const T *hy, const T *hz, const int c1[3], const int c2[3],
                    bool flip, int fa, int irup, hipStream_t s) {
    hipLaunchKernelGGL((k_init_vc<T>), dim3(blocks_of(g.nm)), dim3(256), 0, s, vc, g, hx, hy, hz, i3(c1), i3(c2),
                       flip ? 1 : 0, fa, irup);
}
template <class T>
void launch_init_cells(const Grid<T> &g, const MatInit<T> &m, const T *rho, const T *vp, const T *vs, T *gam, T *rc,
                       T *s1, T *s2, T *lam, T *mu, T *nu, hipStream_t s) {
    hipLaunchKernelGGL((k_init_cells<T>), dim3(blocks_of(g.nm)), dim3(256), 0, s, g, m, rho, vp, vs, gam, rc, s1, s2,
                       lam, mu, nu);
}
template <class T>
void launch_init_minmax(const Grid<T> &g, const MatArrays<T> &a, const int lo[3], const int hi[3], T *part, int nblk,
                        hipStream_t s) {
    hipLaunchKernelGGL((k_init_minmax<T>), dim3(nblk), dim3(256), 0, s, g, a, i3(lo), i3(hi), part);
}
template <class T>
void launch_init_yy(const Grid<T> &g, T *yy, const T *lam, const T *mu, T sc, hipStream_t s) {
    hipLaunchKernelGGL((k_init_yy<T>), dim3(blocks_of(g.nm)), dim3(256), 0, s, g, yy, lam, mu, sc);
}
template <class T>
void launch_init_scale(const Grid<T> &g, T *out, const T *a, const T *b, T c, hipStream_t s) {
    hipLaunchKernelGGL((k_init_scale<T>), dim3(blocks_of(g.nm)), dim3(256), 0, s, g, out, a, b, c);
}
template <class T>
void launch_init_nodeavg(const Grid<T> &g, T *out, const T *in, const int n1[3], const int n2[3], bool inv,
                         hipStream_t s) {
    hipLaunchKernelGGL((k_init_nodeavg<T>), dim3(blocks_of(g.nm)), dim3(256), 0, s, g, out, in, i3(n1), i3(n2),
                       inv ? 1 : 0);
}
template <class T>
void launch_init_recip(const Grid<T> &g, T *a, hipStream_t s) {
    hipLaunchKernelGGL((k_init_recip<T>), dim3(blocks_of(g.nm)), dim3(256), 0, s, g, a);
}
template <class T>
void launch_init_halo(const Grid<T> &g, T *f, T r, const int i1[3], const int i2[3], int pa, int idx, hipStream_t s) {
    hipLaunchKernelGGL((k_init_halo<T>), dim3(blocks_of(g.nm)), dim3(256), 0, s, g, f, r, i3(i1), i3(i2), pa, idx);
}

#define SORD_INIT_INST(T)                                                                                          \
    template void launch_init_vc<T>(T *, const Grid<T> &, const T *, const T *, const T *, const int *, const int *, \
                                    bool, int, int, hipStream_t);                                                  \
    template void launch_init_cells<T>(const Grid<T> &, const MatInit<T> &, const T *, const T *, const T *, T *,  \
                                       T *, T *, T *, T *, T *, T *, hipStream_t);                                 \
    template void launch_init_minmax<T>(const Grid<T> &, const MatArrays<T> &, const int *, const int *, T *, int, \
                                        hipStream_t);                                                              \
    template void launch_init_yy<T>(const Grid<T> &, T *, const T *, const T *, T, hipStream_t);                   \
    template void launch_init_scale<T>(const Grid<T> &, T *, const T *, const T *, T, hipStream_t);                \
    template void launch_init_nodeavg<T>(const Grid<T> &, T *, const T *, const int *, const int *, bool,          \
                                         hipStream_t);                                                             \
    template void launch_init_recip<T>(const Grid<T> &, T *, hipStream_t);                                         \
    template void launch_init_halo<T>(const Grid<T> &, T *, T, const int *, const int *, int, int, hipStream_t);
SORD_INIT_INST(float)
SORD_INIT_INST(double)

}  // namespace sord
