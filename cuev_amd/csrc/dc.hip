// dc.hip — Cuppen divide and conquer for the symmetric tridiagonal eigenproblem
// with the eigenvector matrix resident in device memory.
//
// Algorithm (the published one that LAPACK's ?stedc and the reference's
// tridi.cu:241-418 both follow):
//  * scale T by its largest entry; tear it recursively into halves, subtracting
//    |e| at each cut from the two neighbouring diagonal entries (Cuppen 1981);
//  * leaves (<= leaf size) by implicit QL/QR (host ?steqr, one thread per leaf);
//  * merge T1 (+) T2 + rho v v^T bottom up. Per merge:
//      - z = [last row of Q1, first row of Q2]; rho < 0 flips z2; z /= sqrt 2,
//        rho <- 2|rho| (LAPACK dlaed2's normalisation);
//      - deflation (Dongarra-Sorensen 1987): entries with rho|z_j| <= tol, and
//        close pairs of poles, removed by a Givens rotation when the coupling
//        it drops, |t c s|, is <= tol; tol = 8 eps max(|d|max, |z|max);
//      - the k remaining poles: secular roots lambda_j of
//        1/rho + sum_i z_i^2 / (d_i - lambda) = 0, each as (origin pole, tau)
//        so d_i - lambda_j = (d_i - d_origin) - tau stays accurate;
//      - the Gu-Eisenstat (1995) vector zhat_i^2 =
//        (lambda_i - d_i)/rho * prod_{j != i} (lambda_j - d_i)/(d_j - d_i), which
//        makes the k x k eigenvector matrix S_ij = zhat_i / (d_i - lambda_j)
//        (columns normalised) orthogonal to working accuracy;
//      - Q_new = [Q1 0; 0 Q2] S as two GEMMs over the non-zero row blocks
//        (columns grouped: top-only, dense, bottom-only, as dlaed2 groups them);
//      - results placed in ascending order of eigenvalue, deflated columns copied.
//
// Split of the work: the host does only the O(N log N) deflation bookkeeping
// of each merge (all merges of one tree level in parallel threads) and the
// leaves; the device does the rotations, column gathers, secular roots,
// zhat, S, the GEMMs and the placement. One host-device round trip per tree
// level (z and the current eigenvalues down, the merge descriptors up).
#include "dc.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "lapack.hpp"
#include "profile.hpp"
#include "util.hpp"

namespace cuev {

namespace {

// One merge of a tree level, as the device kernels see it.
struct MergeDesc {
    i64 chunk;      // element offset of this merge's Qc / S / Qn blocks (sum of N^2 before it)
    int off, N, n1; // rows and columns [off, off + N); the cut after n1
    int k, k1, k2;  // non-deflated count; of them k1 top-only, k2 dense, k - k1 - k2 bottom-only
    int ks;         // leading dimension of S (k, or k + 1 to make it even)
    int zb;         // 1: Qc's bottom block starts with a zero column (see the merge GEMMs)
    int rot_begin, rot_end; // Givens rotations of the deflation
    int root_begin;         // flat index of the first non-deflated pole
    int defl_begin;         // flat index of the first deflated entry
    double rho;             // 2|e_cut| (the normalised coupling)
};

struct CopyJob {
    i64 src, dst;
    int len, zero; // zero != 0: write len zeros at dst (src unused)
};

struct LeafDesc {
    int off, sz;
    i64 soff; // element offset of the leaf's eigenvectors in the staging block
};

template <typename T> struct Eps;
template <> struct Eps<double> {
    static constexpr double v = 0.5 * std::numeric_limits<double>::epsilon();
};
template <> struct Eps<float> {
    static constexpr float v = 0.5f * std::numeric_limits<float>::epsilon();
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// leaf eigenvectors (staged column-major per leaf) into the diagonal blocks of Q
template <typename T>
__global__ __launch_bounds__(256) void dc_leaf_kernel(T *Q, i64 ldq, T *dval, const T *stage,
                                                      const T *vals, const int *leaf_of,
                                                      const LeafDesc *lv) {
    const int g = blockIdx.x;
    const LeafDesc L = lv[leaf_of[g]];
    const int c = g - L.off;
    const T *src = stage + L.soff + (i64)c * L.sz;
    T *dst = Q + L.off + (i64)g * ldq;
    for (int r = threadIdx.x; r < L.sz; r += 256) dst[r] = src[r];
    if (threadIdx.x == 0) dval[g] = vals[g];
}

// z of each merge: last row of Q1 for the first n1 columns, first row of Q2 after
template <typename T>
__global__ __launch_bounds__(256) void dc_z_kernel(const T *Q, i64 ldq, const MergeDesc *md, T *z) {
    const MergeDesc m = md[blockIdx.y];
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= m.N) return;
    const int row = m.off + (c < m.n1 ? m.n1 - 1 : m.n1);
    z[m.off + c] = Q[row + (i64)(m.off + c) * ldq];
}

// deflation rotations, in order, on the rows of each merge (one thread per row)
template <typename T>
__global__ __launch_bounds__(256) void dc_rot_kernel(T *Q, i64 ldq, const MergeDesc *md,
                                                     const int2 *ab, const T *rc, const T *rs) {
    const MergeDesc m = md[blockIdx.y];
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (m.rot_begin == m.rot_end || r >= m.N) return;
    T *q = Q + m.off + r;
    for (int t = m.rot_begin; t < m.rot_end; ++t) {
        const int2 p = ab[t];
        const T c = rc[t], s = rs[t];
        const T x = q[(i64)p.x * ldq], y = q[(i64)p.y * ldq];
        q[(i64)p.x * ldq] = c * x + s * y;
        q[(i64)p.y * ldq] = c * y - s * x;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void dc_copy_kernel(const T *src, T *dst, const CopyJob *jobs) {
    const CopyJob j = jobs[blockIdx.x];
    for (int r = threadIdx.x; r < j.len; r += 256) dst[j.dst + r] = j.zero ? T(0) : src[j.src + r];
}

// Secular roots, one wave per root. Root j (0-based) of merge m lies in
// (dl_j, dl_{j+1}) (j < k-1) or (dl_k, dl_k + rho sum w^2] (the last one). It
// is found as tau relative to the nearer pole (origin), so every difference
// d_i - lambda = (d_i - d_o) - tau keeps its relative accuracy. Iteration: a
// two-pole rational model of the secular function fitted to value and
// derivative at the current iterate (Bunch, Nielsen and Sorensen 1978), its
// root accepted when inside the bracket, bisection otherwise; stop when |f|
// is below its rounding-error bound or the bracket is a few ulps wide.
template <typename T>
__global__ __launch_bounds__(256) void dc_secular_kernel(const MergeDesc *md, const int *rmerge,
                                                         const T *dl, const T *w, int nroots,
                                                         int *orig, T *tau, T *lam) {
    const int root = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (root >= nroots) return; // uniform per wave
    const MergeDesc m = md[rmerge[root]];
    const int base = m.root_begin, k = m.k, j = root - base;
    const T *D = dl + base;
    const T *W = w + base;
    const T rho = (T)m.rho, rinv = T(1) / rho;
    const T eps = Eps<T>::v;

    int o;
    T lo, hi;
    if (j < k - 1) {
        const T gap = D[j + 1] - D[j];
        const T h = gap * T(0.5);
        T sm = 0;
        for (int i = lane; i < k; i += 64) sm += W[i] * W[i] / ((D[i] - D[j]) - h);
        sm = wave_sum(sm);
        if (rinv + sm >= T(0)) {
            o = j;
            lo = 0;
            hi = h;
        } else {
            o = j + 1;
            lo = -h;
            hi = 0;
        }
    } else {
        T sw = 0;
        for (int i = lane; i < k; i += 64) sw += W[i] * W[i];
        sw = wave_sum(sw);
        o = j;
        lo = 0;
        hi = rho * sw;
    }
    const T Do = D[o];
    const T Dj = D[j] - Do;
    const T Dj1 = (j < k - 1) ? D[j + 1] - Do : T(0);
    T t = lo + (hi - lo) * T(0.5);
    for (int it = 0; it < 100; ++it) {
        T psi = 0, dpsi = 0, phi = 0, dphi = 0;
        for (int i = lane; i < k; i += 64) {
            const T r = T(1) / ((D[i] - Do) - t); // one division per term
            const T q = W[i] * W[i] * r;
            if (i <= j) {
                psi += q;
                dpsi += q * r;
            } else {
                phi += q;
                dphi += q * r;
            }
        }
        psi = wave_sum(psi);
        dpsi = wave_sum(dpsi);
        phi = wave_sum(phi);
        dphi = wave_sum(dphi);
        const T f = rinv + psi + phi;
        const T err = eps * (T(8) * (rinv + fabs(psi) + phi) + fabs(t) * (dpsi + dphi));
        if (fabs(f) <= err) break;
        if (f < T(0)) lo = t;
        else hi = t;
        if (hi - lo <= T(2) * eps * fmax(fabs(lo), fabs(hi))) break;
        // model: c + b1/(Dj - x) + b2/(Dj1 - x) = 0
        const T dj = Dj - t;
        const T b1 = dpsi * dj * dj;
        const T a1 = psi - b1 / dj;
        T tn = lo - T(1); // marks "no candidate"
        if (j < k - 1) {
            const T dj1 = Dj1 - t;
            const T b2 = dphi * dj1 * dj1;
            const T a2 = phi - b2 / dj1;
            const T c = rinv + a1 + a2;
            const T qa = c, qb = -(c * (Dj + Dj1) + b1 + b2), qc = c * Dj * Dj1 + b1 * Dj1 + b2 * Dj;
            if (qa == T(0)) {
                if (qb != T(0)) tn = -qc / qb;
            } else {
                const T disc = qb * qb - T(4) * qa * qc;
                if (disc >= T(0)) {
                    const T sq = sqrt(disc);
                    const T qq = T(-0.5) * (qb + copysign(sq, qb));
                    const T x1 = qq / qa;
                    const T x2 = (qq != T(0)) ? qc / qq : x1;
                    tn = (x1 > lo && x1 < hi) ? x1 : x2;
                }
            }
        } else {
            const T c = rinv + a1 + phi;
            if (c > T(0)) tn = Dj + b1 / c;
        }
        if (!(tn > lo && tn < hi)) {
            tn = lo + (hi - lo) * T(0.5);
            if (!(tn > lo && tn < hi)) break;
        }
        if (tn == t) break;
        t = tn;
    }
    if (lane == 0) {
        orig[root] = o;
        tau[root] = t;
        lam[root] = Do + t;
    }
}

// Gu-Eisenstat zhat, one wave per pole: lane l multiplies the factors j = l, l + 64, ...
// and the 64 partial products are multiplied in a fixed butterfly order
template <typename T>
__global__ __launch_bounds__(256) void dc_zhat_kernel(const MergeDesc *md, const int *rmerge,
                                                      const T *dl, const T *w, const int *orig,
                                                      const T *tau, T *zh, int nroots) {
    const int gi = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (gi >= nroots) return; // uniform per wave
    const MergeDesc m = md[rmerge[gi]];
    const int base = m.root_begin, k = m.k, i = gi - base;
    const T *D = dl + base;
    const int *O = orig + base;
    const T *TA = tau + base;
    const T di = D[i];
    const T rho = (T)m.rho;
    T p = 1;
    for (int j = lane; j < k; j += 64) {
        const T lmd = -((di - D[O[j]]) - TA[j]); // lambda_j - d_i
        p *= (j == i) ? lmd / rho : lmd / (D[j] - di);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) p *= __shfl_xor(p, off, 64);
    if (lane == 0) zh[gi] = copysign(sqrt(fmax(p, T(0))), w[gi]);
}

// S columns: S[p, j] = zhat_i / (d_i - lambda_j) normalised, p the grouped
// position of pole i (t2i maps positions to poles); one block per root
template <typename T>
__global__ __launch_bounds__(256) void dc_s_kernel(const MergeDesc *md, const int *rmerge,
                                                   const T *dl, const T *zh, const int *t2i,
                                                   const int *orig, const T *tau, T *S) {
    __shared__ T red[4];
    const int gj = blockIdx.x;
    const MergeDesc m = md[rmerge[gj]];
    const int base = m.root_begin, k = m.k, j = gj - base;
    const T *D = dl + base;
    const T *Z = zh + base;
    const int *P = t2i + base;
    const T dor = D[orig[gj]], t = tau[gj];
    T ss = 0;
    for (int p = threadIdx.x; p < k; p += 256) {
        const int i = P[p];
        const T v = Z[i] / ((D[i] - dor) - t);
        ss += v * v;
    }
    ss = wave_sum(ss);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
    __syncthreads();
    const T inv = T(1) / sqrt(red[0] + red[1] + red[2] + red[3]);
    T *col = S + m.chunk + (i64)j * m.ks;
    for (int p = threadIdx.x; p < k; p += 256) {
        const int i = P[p];
        col[p] = (Z[i] / ((D[i] - dor) - t)) * inv;
    }
}

template <typename T> __device__ __forceinline__ int upper_bound_dev(const T *a, int n, T v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
template <typename T> __device__ __forceinline__ int lower_bound_dev(const T *a, int n, T v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// merged columns into Q in ascending order of eigenvalue: roots (Qn columns
// [0, k)) and deflated columns (Qn columns [k, N)) interleaved by value
template <typename T>
__global__ __launch_bounds__(256) void dc_place_kernel(T *Q, i64 ldq, T *dval, const MergeDesc *md,
                                                       const T *Qn, const T *lam, const T *dv) {
    const MergeDesc m = md[blockIdx.y];
    const int c = blockIdx.x;
    if (c >= m.N) return;
    const int kd = m.N - m.k;
    T val;
    int pos;
    if (c < m.k) {
        val = lam[m.root_begin + c];
        pos = c + upper_bound_dev(dv + m.defl_begin, kd, val);
    } else {
        const int t = c - m.k;
        val = dv[m.defl_begin + t];
        pos = t + lower_bound_dev(lam + m.root_begin, m.k, val);
    }
    const T *src = Qn + m.chunk + (i64)c * m.N;
    T *dst = Q + m.off + (i64)(m.off + pos) * ldq;
    for (int r = threadIdx.x; r < m.N; r += 256) dst[r] = src[r];
    if (threadIdx.x == 0) dval[m.off + pos] = val;
}

template <typename T>
__global__ void dc_scale_kernel(const T *src, T *dst, int n, T s) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i] * s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

template <typename F> void parallel_for(int count, F &&f) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::min({count, hw, 16});
    if (nt <= 1) {
        for (int i = 0; i < count; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1)) f(i);
        });
    for (auto &x : th) x.join();
}

struct Node {
    int off, N, n1, depth;
};

void build_tree(int off, int N, int depth, int leaf, std::vector<Node> &merges,
                std::vector<std::pair<int, int>> &leaves, int &maxdepth) {
    if (N <= leaf) {
        leaves.emplace_back(off, N);
        return;
    }
    const int n1 = N / 2;
    merges.push_back({off, N, n1, depth});
    maxdepth = std::max(maxdepth, depth);
    build_tree(off, n1, depth + 1, leaf, merges, leaves, maxdepth);
    build_tree(off + n1, N - n1, depth + 1, leaf, merges, leaves, maxdepth);
}

// Host result of one merge's deflation.
template <typename T> struct MergeHost {
    int k = 0, k1 = 0, k2 = 0;
    double rho = 0;
    std::vector<T> dl, w;              // k poles ascending, weights
    std::vector<int> t2i;              // grouped position -> pole index
    std::vector<int> gcol;             // grouped position -> absolute source column
    std::vector<int2> rot;             // absolute column pairs
    std::vector<T> rc, rs;             // rotation cosines / sines
    std::vector<std::pair<T, int>> defl; // (value, absolute column), ascending
};

// Deflation of one merge (the rules of LAPACK dlaed2, restated).
// d, z: N values of the merge (d ascending within each half).
template <typename T>
void deflate(int off, int N, int n1, T e_cut, const T *d_in, const T *z_in, MergeHost<T> &h) {
    std::vector<T> d(d_in, d_in + N), z(z_in, z_in + N);
    if (e_cut < T(0))
        for (int j = n1; j < N; ++j) z[j] = -z[j];
    const T r2 = T(1) / std::sqrt(T(2));
    for (int j = 0; j < N; ++j) z[j] *= r2;
    const T rho = std::fabs(T(2) * e_cut);
    h.rho = (double)rho;
    std::vector<int> typ(N);
    for (int j = 0; j < N; ++j) typ[j] = j < n1 ? 1 : 3;
    // ascending order of the two sorted halves
    std::vector<int> order(N);
    {
        int a = 0, b = n1, t = 0;
        while (a < n1 && b < N) order[t++] = (d[b] < d[a]) ? b++ : a++;
        while (a < n1) order[t++] = a++;
        while (b < N) order[t++] = b++;
    }
    T dmax = 0, zmax = 0;
    for (int j = 0; j < N; ++j) {
        dmax = std::max(dmax, std::fabs(d[j]));
        zmax = std::max(zmax, std::fabs(z[j]));
    }
    const T tol = T(8) * Eps<T>::v * std::max(dmax, zmax);
    std::vector<int> nd;
    nd.reserve(N);
    h.defl.clear();
    h.rot.clear();
    h.rc.clear();
    h.rs.clear();
    int pj = -1;
    for (int idx = 0; idx < N; ++idx) {
        const int nj = order[idx];
        if (rho * std::fabs(z[nj]) <= tol) {
            h.defl.emplace_back(d[nj], off + nj);
            continue;
        }
        if (pj < 0) {
            pj = nj;
            continue;
        }
        T s = z[pj], c = z[nj];
        const T tau = std::hypot(c, s);
        const T t = d[nj] - d[pj];
        c /= tau;
        s = -s / tau;
        if (std::fabs(t * c * s) <= tol) {
            z[nj] = tau;
            z[pj] = 0;
            if (typ[nj] != typ[pj]) typ[nj] = 2;
            typ[pj] = 4;
            h.rot.push_back(make_int2(off + pj, off + nj));
            h.rc.push_back(c);
            h.rs.push_back(s);
            const T tp = d[pj] * c * c + d[nj] * s * s;
            d[nj] = d[pj] * s * s + d[nj] * c * c;
            d[pj] = tp;
            h.defl.emplace_back(d[pj], off + pj);
            pj = nj;
        } else {
            nd.push_back(pj);
            pj = nj;
        }
    }
    if (pj >= 0) nd.push_back(pj);
    std::sort(h.defl.begin(), h.defl.end(),
              [](const std::pair<T, int> &a, const std::pair<T, int> &b) { return a.first < b.first; });
    const int k = (int)nd.size();
    h.k = k;
    h.dl.resize(k);
    h.w.resize(k);
    for (int i = 0; i < k; ++i) {
        h.dl[i] = d[nd[i]];
        h.w[i] = z[nd[i]];
    }
    // grouped order: top-only, dense, bottom-only (poles ascending within each)
    h.t2i.clear();
    h.gcol.clear();
    h.k1 = h.k2 = 0;
    for (int ty : {1, 2, 3})
        for (int i = 0; i < k; ++i)
            if (typ[nd[i]] == ty) {
                h.t2i.push_back(i);
                h.gcol.push_back(off + nd[i]);
                if (ty == 1) ++h.k1;
                if (ty == 2) ++h.k2;
            }
}

// Fixed-capacity layout of the per-level upload block (identical on host and device).
struct UploadLayout {
    // [md, leaf_of) is uploaded whole once per level; md0 (the level's geometry,
    // uploaded before the deflation) and the leaf tables lie after it
    size_t md, rot_ab, rot_c, rot_s, dl, w, t2i, rmerge, gjob, djob, dv, leaf_of, leaves, md0, total;
};

template <typename T> UploadLayout make_layout(int n) {
    UploadLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += (bytes + 255) & ~size_t(255);
        return at;
    };
    const size_t nn = (size_t)std::max(n, 1);
    L.md = take(sizeof(MergeDesc) * nn);
    L.rot_ab = take(sizeof(int2) * nn);
    L.rot_c = take(sizeof(T) * nn);
    L.rot_s = take(sizeof(T) * nn);
    L.dl = take(sizeof(T) * nn);
    L.w = take(sizeof(T) * nn);
    L.t2i = take(sizeof(int) * nn);
    L.rmerge = take(sizeof(int) * nn);
    L.gjob = take(sizeof(CopyJob) * 2 * nn);
    L.djob = take(sizeof(CopyJob) * nn);
    L.dv = take(sizeof(T) * nn);
    L.leaf_of = take(sizeof(int) * nn);
    L.leaves = take(sizeof(LeafDesc) * nn);
    L.md0 = take(sizeof(MergeDesc) * nn);
    L.total = o;
    return L;
}

} // namespace

template <typename T> struct DcState {
    int n = 0;
    UploadLayout L{};
    char *h_up = nullptr;  // pinned upload block
    char *d_up = nullptr;  // device copy
    T *h_zd = nullptr;     // pinned: z (n) then current eigenvalues (n)
    T *d_z = nullptr;      // device z (n)
    T *d_val = nullptr;    // device current eigenvalues (n)
    int *d_orig = nullptr; // secular results (n each)
    T *d_tau = nullptr, *d_lam = nullptr, *d_zh = nullptr;
    void *pool = nullptr;
    std::vector<T> leaf_stage; // host leaf eigenvectors, then eigenvalues
};

template <typename T> int dc_state_create(DcState<T> *&st, int n) {
    st = new DcState<T>();
    st->n = n;
    st->L = make_layout<T>(n);
    const size_t nn = (size_t)std::max(n, 1);
    size_t dev = st->L.total + sizeof(T) * nn * 6 + sizeof(int) * nn + 4096;
    if (hipMalloc(&st->pool, dev) != hipSuccess) {
        delete st;
        st = nullptr;
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "dc: hipMalloc(%zu) failed", dev);
    }
    char *p = static_cast<char *>(st->pool);
    auto carve = [&](size_t bytes) {
        char *at = p;
        p += (bytes + 255) & ~size_t(255);
        return at;
    };
    st->d_up = carve(st->L.total);
    st->d_z = reinterpret_cast<T *>(carve(sizeof(T) * nn));
    st->d_val = reinterpret_cast<T *>(carve(sizeof(T) * nn));
    st->d_tau = reinterpret_cast<T *>(carve(sizeof(T) * nn));
    st->d_lam = reinterpret_cast<T *>(carve(sizeof(T) * nn));
    st->d_zh = reinterpret_cast<T *>(carve(sizeof(T) * nn));
    st->d_orig = reinterpret_cast<int *>(carve(sizeof(int) * nn));
    if (hipHostMalloc(reinterpret_cast<void **>(&st->h_up), st->L.total, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&st->h_zd), 2 * sizeof(T) * nn,
                      hipHostMallocDefault) != hipSuccess) {
        dc_state_destroy(st);
        st = nullptr;
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "dc: hipHostMalloc failed");
    }
    return CUEV_STATUS_SUCCESS;
}

template <typename T> void dc_state_destroy(DcState<T> *st) {
    if (!st) return;
    if (st->pool) (void)hipFree(st->pool);
    if (st->h_up) (void)hipHostFree(st->h_up);
    if (st->h_zd) (void)hipHostFree(st->h_zd);
    delete st;
}

template <typename T>
int dc_solve(DcState<T> &st, hipStream_t s, int n, const T *d, const T *e, T *Q, i64 ldq, T *eval,
             T *Qc, T *S, T *Qn, Scratch gs, int leaf, hipEvent_t gate, int gate_from) {
    if (n <= 0) return CUEV_STATUS_SUCCESS;
    if (n > st.n) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "dc: state sized for %d, got n=%d", st.n, n);
    if (ldq < n) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "dc: ldq < n");
    if (leaf <= 0) leaf = 64;
    leaf = std::max(leaf, 2);
    const LapackFns *lp = lapack();
    if (!lp || (sizeof(T) == 8 ? !lp->dsteqr : !lp->ssteqr))
        CUEV_FAIL(CUEV_STATUS_LAPACK_UNAVAILABLE, "dc: host ?steqr not found");

    // scale
    T orgnrm = 0;
    for (int i = 0; i < n; ++i) orgnrm = std::max(orgnrm, (T)std::fabs(d[i]));
    for (int i = 0; i + 1 < n; ++i) orgnrm = std::max(orgnrm, (T)std::fabs(e[i]));
    CUEV_TRY(fill2d<T>(s, Q, ldq, n, n, T(0), T(1)));
    if (orgnrm == T(0)) {
        CUEV_TRY((fill2d<T>(s, eval, n, n, 1, T(0), T(0))));
        CUEV_HIP(hipStreamSynchronize(s));
        return CUEV_STATUS_SUCCESS;
    }
    std::vector<T> ds(n), es(std::max(n - 1, 0));
    for (int i = 0; i < n; ++i) ds[i] = d[i] / orgnrm;
    for (int i = 0; i + 1 < n; ++i) es[i] = e[i] / orgnrm;

    std::vector<Node> merges;
    std::vector<std::pair<int, int>> leaves;
    int maxdepth = -1;
    build_tree(0, n, 0, leaf, merges, leaves, maxdepth);
    for (const Node &m : merges) {
        const T r = std::fabs(es[m.off + m.n1 - 1]);
        ds[m.off + m.n1 - 1] -= r;
        ds[m.off + m.n1] -= r;
    }

    // ---- leaves on the host, in parallel
    const UploadLayout &L = st.L;
    char *up = st.h_up;
    LeafDesc *lv = reinterpret_cast<LeafDesc *>(up + L.leaves);
    int *leaf_of = reinterpret_cast<int *>(up + L.leaf_of);
    i64 soff = 0;
    for (size_t i = 0; i < leaves.size(); ++i) {
        lv[i] = {leaves[i].first, leaves[i].second, soff};
        for (int c = 0; c < leaves[i].second; ++c) leaf_of[leaves[i].first + c] = (int)i;
        soff += (i64)leaves[i].second * leaves[i].second;
    }
    const i64 vals_off = soff; // sum of leaf sizes squared <= n^2: fits Qn
    st.leaf_stage.resize((size_t)(vals_off + n));
    T *stage = st.leaf_stage.data();
    std::atomic<int> leaf_info{0};
    {
        prof::Scope sc(s, "dc.leaves");
        parallel_for((int)leaves.size(), [&](int i) {
            const int off = lv[i].off, sz = lv[i].sz;
            T *z = stage + lv[i].soff;
            T *dv = stage + vals_off + off;
            std::vector<T> ee(std::max(sz - 1, 1)), work(std::max(2 * sz - 2, 1));
            for (int j = 0; j < sz; ++j) dv[j] = ds[off + j];
            for (int j = 0; j + 1 < sz; ++j) ee[j] = es[off + j];
            const char compz = 'I';
            int nn = sz, ldz = sz, info = 0;
            if constexpr (sizeof(T) == 8)
                lp->dsteqr(&compz, &nn, (double *)dv, (double *)ee.data(), (double *)z, &ldz,
                           (double *)work.data(), &info);
            else
                lp->ssteqr(&compz, &nn, (float *)dv, (float *)ee.data(), (float *)z, &ldz,
                           (float *)work.data(), &info);
            if (info != 0) leaf_info.store(info);
        });
    }
    if (leaf_info.load() != 0)
        CUEV_FAIL(CUEV_STATUS_LAPACK_ERROR, "dc: ?steqr on a leaf returned %d", leaf_info.load());
    // leaf eigenvectors through Qn (free until the first merge), eigenvalues through d_z
    CUEV_HIP(hipMemcpyAsync(Qn, stage, sizeof(T) * (size_t)vals_off, hipMemcpyHostToDevice, s));
    CUEV_HIP(hipMemcpyAsync(st.d_z, stage + vals_off, sizeof(T) * (size_t)n, hipMemcpyHostToDevice, s));
    CUEV_HIP(hipMemcpyAsync(st.d_up + L.leaf_of, up + L.leaf_of, sizeof(int) * (size_t)n,
                            hipMemcpyHostToDevice, s));
    CUEV_HIP(hipMemcpyAsync(st.d_up + L.leaves, up + L.leaves, sizeof(LeafDesc) * leaves.size(),
                            hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(dc_leaf_kernel<T>, dim3((unsigned)n), dim3(256), 0, s, Q, ldq, st.d_val, Qn,
                       st.d_z, reinterpret_cast<const int *>(st.d_up + L.leaf_of),
                       reinterpret_cast<const LeafDesc *>(st.d_up + L.leaves));
    CUEV_LAUNCH_CHECK();

    // ---- merges, deepest level first
    MergeDesc *md = reinterpret_cast<MergeDesc *>(up + L.md);
    MergeDesc *md0 = reinterpret_cast<MergeDesc *>(up + L.md0);
    const MergeDesc *d_md = reinterpret_cast<const MergeDesc *>(st.d_up + L.md);
    const MergeDesc *d_md0 = reinterpret_cast<const MergeDesc *>(st.d_up + L.md0);
    std::vector<MergeHost<T>> mh;
    T *hz = st.h_zd, *hval = st.h_zd + n;
    for (int depth = maxdepth; depth >= 0; --depth) {
        std::vector<const Node *> lev;
        for (const Node &m : merges)
            if (m.depth == depth) lev.push_back(&m);
        const int nm = (int)lev.size();
        if (nm == 0) continue;
        // The previous level's upload of [md, leaf_of) may still be in flight: the
        // geometry goes through md0, and `up` below md0 is rewritten only after
        // the synchronisation that follows the z download.
        int maxN = 0;
        for (int i = 0; i < nm; ++i) {
            md0[i] = MergeDesc{};
            md0[i].off = lev[i]->off;
            md0[i].N = lev[i]->N;
            md0[i].n1 = lev[i]->n1;
            maxN = std::max(maxN, lev[i]->N);
        }
        if (gate && maxN >= gate_from) {
            // from here on the merges have the GPU to themselves: full-size tiles
            CUEV_HIP(hipStreamWaitEvent(s, gate, 0));
            gate = nullptr;
            gs.small_tiles = 0;
        }
        // z and the current eigenvalues to the host
        CUEV_HIP(hipMemcpyAsync(st.d_up + L.md0, md0, sizeof(MergeDesc) * nm, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(dc_z_kernel<T>, dim3((unsigned)ceil_div(maxN, 256), (unsigned)nm), dim3(256),
                           0, s, Q, ldq, d_md0, st.d_z);
        CUEV_LAUNCH_CHECK();
        CUEV_HIP(hipMemcpyAsync(hz, st.d_z, sizeof(T) * n, hipMemcpyDeviceToHost, s));
        CUEV_HIP(hipMemcpyAsync(hval, st.d_val, sizeof(T) * n, hipMemcpyDeviceToHost, s));
        CUEV_HIP(hipStreamSynchronize(s));

        mh.resize(nm);
        parallel_for(nm, [&](int i) {
            const Node &m = *lev[i];
            deflate<T>(m.off, m.N, m.n1, es[m.off + m.n1 - 1], hval + m.off, hz + m.off, mh[i]);
        });

        // pack the level
        int2 *rot_ab = reinterpret_cast<int2 *>(up + L.rot_ab);
        T *rot_c = reinterpret_cast<T *>(up + L.rot_c);
        T *rot_s = reinterpret_cast<T *>(up + L.rot_s);
        T *dl = reinterpret_cast<T *>(up + L.dl);
        T *w = reinterpret_cast<T *>(up + L.w);
        int *t2i = reinterpret_cast<int *>(up + L.t2i);
        int *rmerge = reinterpret_cast<int *>(up + L.rmerge);
        CopyJob *gjob = reinterpret_cast<CopyJob *>(up + L.gjob);
        CopyJob *djob = reinterpret_cast<CopyJob *>(up + L.djob);
        T *dv = reinterpret_cast<T *>(up + L.dv);
        int nrot = 0, nroot = 0, ndefl = 0, ng = 0;
        i64 chunk = 0;
        for (int i = 0; i < nm; ++i) {
            const MergeHost<T> &h = mh[i];
            MergeDesc &m = md[i];
            m = md0[i];
            const int N = m.N, n1 = m.n1, n2 = N - n1, k = h.k;
            m.chunk = chunk;
            m.k = k;
            m.k1 = h.k1;
            m.k2 = h.k2;
            m.rho = h.rho;
            m.rot_begin = nrot;
            for (size_t r = 0; r < h.rot.size(); ++r) {
                rot_ab[nrot] = h.rot[r];
                rot_c[nrot] = h.rc[r];
                rot_s[nrot] = h.rs[r];
                ++nrot;
            }
            m.rot_end = nrot;
            m.root_begin = nroot;
            for (int p = 0; p < k; ++p) {
                dl[nroot + p] = h.dl[p];
                w[nroot + p] = h.w[p];
                t2i[nroot + p] = h.t2i[p];
                rmerge[nroot + p] = i;
            }
            const int ktop = h.k1 + h.k2, kbot = k - h.k1;
            // The merge GEMMs read S with an even leading dimension and the bottom block's
            // S rows from an even row, so that both operands can take the engine's 16-byte
            // (LDS-DMA) path: S gets ld k + 1 when k is odd, and when k1 is odd the bottom
            // block of Qc starts with a zero column and its product reads S from row k1 - 1
            // (a top-only row, which the zero column cancels). Each only when it fits the
            // merge's N^2 block.
            m.ks = ((k & 1) && (i64)(k + 1) * k <= (i64)N * N) ? k + 1 : k;
            m.zb = ((h.k1 & 1) && (i64)n1 * ktop + (i64)n2 * (kbot + 1) <= (i64)N * N) ? 1 : 0;
            for (int p = 0; p < ktop; ++p)
                gjob[ng++] = CopyJob{(i64)m.off + (i64)h.gcol[p] * ldq, chunk + (i64)p * n1, n1, 0};
            if (m.zb) gjob[ng++] = CopyJob{0, chunk + (i64)n1 * ktop, n2, 1};
            for (int p = h.k1; p < k; ++p)
                gjob[ng++] = CopyJob{(i64)(m.off + n1) + (i64)h.gcol[p] * ldq,
                                     chunk + (i64)n1 * ktop + (i64)(p - h.k1 + m.zb) * n2, n2, 0};
            m.defl_begin = ndefl;
            for (size_t t = 0; t < h.defl.size(); ++t) {
                dv[ndefl + (int)t] = h.defl[t].first;
                djob[ndefl + (int)t] = CopyJob{(i64)m.off + (i64)h.defl[t].second * ldq,
                                               chunk + (i64)(k + (int)t) * N, N, 0};
            }
            ndefl += (int)h.defl.size();
            nroot += k;
            chunk += (i64)N * N;
        }
        if (ng > 2 * n || nroot > n || ndefl > n || nrot > n)
            CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "dc: level bookkeeping exceeds its capacity");
        CUEV_HIP(hipMemcpyAsync(st.d_up, up, L.leaf_of, hipMemcpyHostToDevice, s));

        const char *du = st.d_up;
        if (nrot > 0) {
            prof::Scope sc(s, "dc.rotate");
            hipLaunchKernelGGL(dc_rot_kernel<T>, dim3((unsigned)ceil_div(maxN, 256), (unsigned)nm),
                               dim3(256), 0, s, Q, ldq, d_md,
                               reinterpret_cast<const int2 *>(du + L.rot_ab),
                               reinterpret_cast<const T *>(du + L.rot_c),
                               reinterpret_cast<const T *>(du + L.rot_s));
            CUEV_LAUNCH_CHECK();
        }
        {
            prof::Scope sc(s, "dc.gather");
            if (ng > 0) {
                hipLaunchKernelGGL(dc_copy_kernel<T>, dim3((unsigned)ng), dim3(256), 0, s, Q, Qc,
                                   reinterpret_cast<const CopyJob *>(du + L.gjob));
                CUEV_LAUNCH_CHECK();
            }
            if (ndefl > 0) {
                hipLaunchKernelGGL(dc_copy_kernel<T>, dim3((unsigned)ndefl), dim3(256), 0, s, Q, Qn,
                                   reinterpret_cast<const CopyJob *>(du + L.djob));
                CUEV_LAUNCH_CHECK();
            }
        }
        const T *d_dl = reinterpret_cast<const T *>(du + L.dl);
        const T *d_w = reinterpret_cast<const T *>(du + L.w);
        const int *d_rm = reinterpret_cast<const int *>(du + L.rmerge);
        if (nroot > 0) {
            {
                prof::Scope sc(s, "dc.secular");
                hipLaunchKernelGGL(dc_secular_kernel<T>, dim3((unsigned)ceil_div(nroot, 4)), dim3(256),
                                   0, s, d_md, d_rm, d_dl, d_w, nroot, st.d_orig, st.d_tau, st.d_lam);
                CUEV_LAUNCH_CHECK();
            }
            {
                prof::Scope sc(s, "dc.zhat_s");
                hipLaunchKernelGGL(dc_zhat_kernel<T>, dim3((unsigned)ceil_div(nroot, 4)), dim3(256),
                                   0, s, d_md, d_rm, d_dl, d_w, st.d_orig, st.d_tau, st.d_zh, nroot);
                CUEV_LAUNCH_CHECK();
                hipLaunchKernelGGL(dc_s_kernel<T>, dim3((unsigned)nroot), dim3(256), 0, s, d_md, d_rm,
                                   d_dl, st.d_zh, reinterpret_cast<const int *>(du + L.t2i),
                                   st.d_orig, st.d_tau, S);
                CUEV_LAUNCH_CHECK();
            }
            double fl = 0;
            for (int i = 0; i < nm; ++i) {
                const MergeDesc &m = md[i];
                fl += 2.0 * m.k * ((double)m.n1 * (m.k1 + m.k2) + (double)(m.N - m.n1) * (m.k - m.k1));
            }
            prof::Scope sc(s, "dc.gemm", fl);
            for (int i = 0; i < nm; ++i) {
                const MergeDesc &m = md[i];
                if (m.k == 0) continue;
                const int N = m.N, n1 = m.n1, n2 = N - n1, k = m.k, ktop = m.k1 + m.k2,
                          kbot = m.k - m.k1;
                const T *qc = Qc + m.chunk;
                const T *sm = S + m.chunk;
                T *qn = Qn + m.chunk;
                CUEV_TRY(gemm<T>(s, 0, 0, n1, k, ktop, T(1), qc, n1, sm, m.ks, T(0), qn, N, gs));
                CUEV_TRY(gemm<T>(s, 0, 0, n2, k, kbot + m.zb, T(1), qc + (i64)n1 * ktop, std::max(n2, 1),
                                 sm + m.k1 - m.zb, m.ks, T(0), qn + n1, N, gs));
            }
        }
        {
            prof::Scope sc(s, "dc.place");
            hipLaunchKernelGGL(dc_place_kernel<T>, dim3((unsigned)maxN, (unsigned)nm), dim3(256), 0, s, Q,
                               ldq, st.d_val, d_md, Qn, st.d_lam, reinterpret_cast<const T *>(du + L.dv));
            CUEV_LAUNCH_CHECK();
        }
    }
    hipLaunchKernelGGL(dc_scale_kernel<T>, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, st.d_val,
                       eval, n, orgnrm);
    CUEV_LAUNCH_CHECK();
    CUEV_HIP(hipStreamSynchronize(s));
    return CUEV_STATUS_SUCCESS;
}

template int dc_state_create<float>(DcState<float> *&, int);
template int dc_state_create<double>(DcState<double> *&, int);
template void dc_state_destroy<float>(DcState<float> *);
template void dc_state_destroy<double>(DcState<double> *);
template int dc_solve<float>(DcState<float> &, hipStream_t, int, const float *, const float *, float *,
                             i64, float *, float *, float *, float *, Scratch, int, hipEvent_t, int);
template int dc_solve<double>(DcState<double> &, hipStream_t, int, const double *, const double *,
                              double *, i64, double *, double *, double *, double *, Scratch, int,
                              hipEvent_t, int);

} // namespace cuev
