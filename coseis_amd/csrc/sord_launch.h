// Host-callable launchers of the SORD kernels (explicitly instantiated for
// float and double in sord_kernels.hip and sord_fused.hip).
#pragma once
#include "sord_dev.h"

namespace sord {
template <class T> void launch_finite_source(const FSrc<T> &S, T tm, const Params<T> &P, bool box, hipStream_t s);
template <class T> void launch_stress(const Params<T> &P, const PointOps<T> &ops, int3 lo, int3 hi, int mode, hipStream_t s);
template <class T> void launch_accel(const Params<T> &P, int3 lo, int3 hi, int mode, hipStream_t s);
template <class T> void launch_time(const Params<T> &P, int3 lo, int3 hi, hipStream_t s);
template <class T> void launch_newton(const Params<T> &P, int3 lo, int3 hi, hipStream_t s);
template <class T> void launch_gather(const T *f, View g, const int lo[3], const int hi[3], const int st[3], const int n[3], T *out, hipStream_t s);
template <class T> void launch_boxop(T *f, View g, const int lo[3], const int hi[3], const int st[3], const int n[3], int op, T val, const T *src, hipStream_t s);
template <class T> void launch_plane_copy(T *f, View g, const int n[3], int ax, int dst, int src, hipStream_t s);
template <class T> void launch_sample(const T *f, View g, const Samples<T> &smp, T *out, hipStream_t s);
template <class T> void launch_point_add(T *f, View g, const int lo[3], const int hi[3], const T xi[3], T val, hipStream_t s);
template <class T> void launch_norm(T *out, T *const a[3], T *const b[3], const Grid<T> &g, const int lo[3], const int hi[3], const int st[3], const int n[3], hipStream_t s);
void launch_copy_words(const void *src, void *dst, long nwords, hipStream_t s);
// device address of the static scratch (SCRATCH_BYTES) on `device`, which must
// be current; nullptr if it cannot be taken
void *static_scratch(int device);
template <class T> int launch_maxnorm(T *const a[3], T *const b[3], const Grid<T> &g, const int lo[3], const int hi[3], T *pval, long *ploc, hipStream_t s);
template <class T> void launch_vbc_plane(const Params<T> &P, int ax, int dst, int src, int mask, const T sg[3], int kind, hipStream_t s);
template <class T> void launch_rupture(const Params<T> &P, const Fault<T> &F, int it, T tm, int fused, hipStream_t s);
template <class T> void launch_fault_time(const Params<T> &P, const Fault<T> &F, T tm, T svtol, hipStream_t s);
template <class T> void launch_fault_svold(const Params<T> &P, const Fault<T> &F, hipStream_t s);
template <class T> void launch_fault_sa(const Params<T> &P, const Fault<T> &F, hipStream_t s);
// device-side initialisation (sord_init.hip)
template <class T> void launch_init_vc(T *vc, const Grid<T> &g, const T *hx, const T *hy, const T *hz, const int c1[3], const int c2[3], bool flip, int fa, int irup, hipStream_t s);
template <class T> void launch_init_cells(const Grid<T> &g, const MatInit<T> &m, const T *rho, const T *vp, const T *vs, T *gam, T *rc, T *s1, T *s2, T *lam, T *mu, T *nu, hipStream_t s);
template <class T> void launch_init_minmax(const Grid<T> &g, const MatArrays<T> &a, const int lo[3], const int hi[3], T *part, int nblk, hipStream_t s);
template <class T> void launch_init_yy(const Grid<T> &g, T *yy, const T *lam, const T *mu, T sc, hipStream_t s);
template <class T> void launch_init_scale(const Grid<T> &g, T *out, const T *a, const T *b, T c, hipStream_t s);
template <class T> void launch_init_nodeavg(const Grid<T> &g, T *out, const T *in, const int n1[3], const int n2[3], bool inv, hipStream_t s);
template <class T> void launch_init_recip(const Grid<T> &g, T *a, hipStream_t s);
template <class T> void launch_init_halo(const Grid<T> &g, T *f, T r, const int i1[3], const int i2[3], int pa, int idx, hipStream_t s);
}  // namespace sord

namespace sord {
// mask -1: interior kernel with `waves` waves per block; 0..7: boundary kernel
// specialised on the damped axes (bit a: axis a)
// e0, e1: when given, the kernel's own start and end times are recorded in them
template <class T> void launch_fused(const Params<T> &P, const PointOps<T> &ops, const Work *work, int nwork, int mask, const T sp[3], int waves, bool fs, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// constants of the fused kernels for one scenario; the batch launch takes a
// device table of scenarios, each work item naming its own (Work.pad1)
template <class T> int boundary_blocks_per_cu();
template <class T> FusedC<T> fused_consts(const Params<T> &P, const T sp[3], const Work *work);
template <class T> void launch_adjoint(const AdjP<T> &A, const Work *work, int nwork, hipStream_t s);
template <class T> void launch_adj_chain(const AdjChain<T> &c, hipStream_t s);
template <class T> void launch_adj_inject(T *const vh[3], const Grid<T> &g, const int *rec, int n, const T *res, hipStream_t s);
template <class T> void launch_adj_gm_end(T *gm, const T *mr, T *const vh[3], const T *const v[3], const Grid<T> &g, const int *node, int n, hipStream_t s);
template <class T> void launch_rec_gather(T *const v[3], const Grid<T> &g, const int *rec, int n, T *out, hipStream_t s);
template <class T> void launch_rec_gather_states(const T *const *states, int nstate, size_t S, const Grid<T> &g, const int *rec, int n, T *out, hipStream_t s);
template <class T> void launch_zero_stats(const StatZero<T> *z, int n, hipStream_t s);
template <class T> void launch_fused_batch(const Scen<T> *sc, const Work *work, int nwork, int mask, bool st, hipStream_t s);
}
