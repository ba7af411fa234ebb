// Class-batched translations as tiled FP64 matrix-core (MFMA) products (gfx950).
//
// Translations whose vectors share the polar angle (a rotation run: e.g. the
// M2L offsets with one |o| and one o_z) share one Wigner-d table; translations
// whose vectors share the length (an axial segment: one |o|) share one coaxial
// table.  Every translation becomes a column of three dense products:
//
//   forward rotation, degree nu:  Y^nu (rows k)  = D^nu (k x m) * X^nu (m x cols)
//   coaxial shift, order k:       Z_k^+ = P_k Y_k^+,  Z_k^- = Q_k Y_k^-            (complex)
//   back rotation, degree n:      W^n  (rows m)  = D^n^T (m x k) * Z^n (k x cols)
//
// The rotations use the folded tables E, O of d^n (half the products' work).
//
// with x_+- = x_a +- x_b and P = A + B, Q = A - B (the [A B; B A] coaxial operator
// decoupled).  Columns are the translations of the class, 4 reals each
// (+.re, +.im, -.re, -.im).
// The gather of the sources with the per-column phase e^{i m phi} is fused into
// the loads of the forward product and the phase e^{-i m phi} into the epilogue
// of the back product, which writes each column as a full expansion W.  A
// separate pass adds the columns of each target in a fixed order (deterministic,
// no atomics).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ifs_internal.h"

namespace ifs {

namespace {

typedef double dv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));

// Matrix-core tile: D(16x16) += A(16x4) B(4x16) with lane l holding A[l%16][l/16] and
// B[l/16][l%16] in both the FP64 (v_mfma_f64_16x16x4_f64) and the FP32
// (v_mfma_f32_16x16x4_f32, FP32 accumulation) form.  Their D layouts differ: value v
// of lane l is D[drow(v, l/16)][l%16] with drow = 4v + l/16 (FP64) or 4(l/16) + v
// (FP32).  The products run in the compute type T; tables and pools stay FP64 in
// memory and are converted where they are staged into LDS, so one set of tables
// serves both.
template <typename T> struct Mx;
template <> struct Mx<double> {
    typedef dv4 v4;
    static constexpr bool F32 = false;
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int v, int lk) { return 4 * v + lk; }
};
template <> struct Mx<float> {
    typedef fv4 v4;
    static constexpr bool F32 = true;
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int v, int lk) { return 4 * lk + v; }
};
template <typename T> struct V2;
template <> struct V2<double> { typedef double2 t; };
template <> struct V2<float> { typedef float2 t; };

__device__ __forceinline__ int fidx(int n, int m) { return n * n - 1 + m + n; }

constexpr int RPAD = 2;  // LDS row padding (doubles)

}  // namespace

// Folded rotation products.  With E, O the folded Wigner tables (k, m >= 0):
//   forward, degree nu:  S = E s,  T = O t,   s_m = u_m + (-1)^m u_{-m} (s_0 = u_0),
//                        t_m = u_m - (-1)^m u_{-m},  u_m = e^{i m phi} x_m
//                        y_k = (S_k + T_k)/2,  y_{-k} = (-1)^k (S_k - T_k)/2
//   back, degree n:      F = E^T s,  G = O^T t,  s_k = z_k + (-1)^k z_{-k} (s_0 = z_0),
//                        t_k = z_k - (-1)^k z_{-k}
//                        w_m = (F_m + G_m)/2,  w_{-m} = (-1)^m (F_m - G_m)/2
// which is half the work of the unfolded products.  A run of polar angle pi - b
// (mirror) uses the tables of b: with sigma = (-1)^{n+k+m},
//   E(pi - b) = sigma E(b),  O(pi - b) = -sigma O(b),
// so the column entries take (-1)^m (forward) or (-1)^k (back) at the LDS store and
// the results (-1)^{n+k} (forward) or (-1)^{n+m} (back), with T and G negated.
// Back product block: FMB = 64 rows of each product x FNB = 64 real columns; wave w
// owns rows 32 (w / 2) .. + 31 and columns 32 (w % 2) .. + 31 of both products, so
// F and G of one entry meet in the same lane; registers bounded for three waves per
// SIMD (128-row blocks fit only two and were slower).  Forward product block: FMF =
// 64 rows of each product x FNF2 = 64 columns; waves 0,1 compute E, waves 2,3 O (32
// columns each), and S and T meet through LDS in the epilogue.  Each row block
// gathers its column panel again, so tall forward blocks pay: 64 rows at three waves
// per SIMD took the forward product from 0.75 to 0.65 s per matvec against 32 rows.
// The narrow column tiles buy occupancy: a tile's operands are loaded one inner step
// ahead, and three waves per SIMD hide their latency better than two.
constexpr int FMB = 64;
constexpr int FMF = 64;
constexpr int FNF2 = 64;
constexpr int FNB = 64;
constexpr int FK = 16;

template <typename TC, typename TW, int RM, int MINB>
__global__ __launch_bounds__(256, MINB) void k_class_rot_fwd(const GroupDesc *__restrict__ desc,
                                                       const int64_t *__restrict__ src_off,
                                                       const int32_t *__restrict__ pidx,
                                                       const double2 *__restrict__ phase,
                                                       const double2 *__restrict__ src_pool,
                                                       const TW *__restrict__ wd, TC *__restrict__ wy) {
    typedef typename Mx<TC>::v4 acc4;
    const GroupDesc C = *desc;
    const int nu = blockIdx.z + 1;
    if (nu > C.Ls) return;
    const int K = C.K;
    const int kk = nu < K ? nu : K;
    const int M = kk + 1;   // rows k = 0..kk
    const int KI = nu + 1;  // inner m = 0..nu
    const int NTOT = C.ncols * C.nrhs;  // columns: translations x right-hand sides
    const int NC = 4 * NTOT;
    const int NCY = 4 * C.ncy * C.nrhs;
    constexpr int NI = RM / 16;  // 16-row MFMA tiles of a wave
    constexpr int NE = RM / 8;   // table entries a thread loads per step (2 RM rows x 16 inner)
    const int row0 = blockIdx.y * RM;
    const int col0 = blockIdx.x * FNF2;
    if (row0 >= M) return;
    // LDS: the table tile As (columns 0..RM-1: E rows, RM..2RM-1: O rows) and the s, t
    // tiles Bs; after the loop the same memory passes T between the waves (X)
    constexpr int AS = FK * (2 * RM + RPAD), BS = 2 * FK * (FNF2 + RPAD), XS = RM * (FNF2 + RPAD);
    __shared__ TC smem[AS + BS > XS ? AS + BS : XS];
    TC(*As)[2 * RM + RPAD] = reinterpret_cast<TC(*)[2 * RM + RPAD]>(smem);
    TC(*Bs)[FK][FNF2 + RPAD] = reinterpret_cast<TC(*)[FK][FNF2 + RPAD]>(smem + AS);
    const TW *E = wd + C.off_d + fold_off(nu, K);
    const TW *O = E + (int64_t)M * KI;
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int g = w >> 1;            // 0: E product, 1: O product
    const int wc = (w & 1) * 32;
    acc4 acc[NI][2];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = acc4{0, 0, 0, 0};
    const int nsrc = C.Ls * (C.Ls + 2);
    // Loads without bounds branches.  A thread's table rows ((tid >> 4) + 16 e of E,
    // then of O) and translation (tid >> 4 of the tile) are clamped into the operands:
    // rows >= M and columns >= NC only feed outputs that are never stored.  Inner
    // entries m >= KI occur in the last step only, which clamps them and zeroes their
    // table entries.
    const int q = tid & 15;
    int roff[NE / 2];  // row offsets, shared by E and O
#pragma unroll
    for (int e = 0; e < NE / 2; ++e) roff[e] = min(row0 + (tid >> 4) + 16 * e, M - 1) * KI;
    const int jt = min((col0 >> 2) + (tid >> 4), NTOT - 1);
    const int tt = jt % C.ncols, rr = jt / C.ncols;  // translation, right-hand side
    const double2 *xs = src_pool + rr * C.src_stride + src_off[C.col_begin + tt] + fidx(nu, 0);  // xs[m]
    const double2 *ps = phase + (int64_t)pidx[C.col_begin + tt] * (2 * C.PL + 1) + C.PL;        // e^{i m phi}
    TW ra[NE];
    double2 rap[1], rbp[1], ram[1], rbm[1], rph[1];
    auto load = [&](int k0, auto tail) {
        int gq = k0 + q;
        if constexpr (decltype(tail)::value) gq = min(gq, KI - 1);
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            TW v = (e < NE / 2 ? E : O)[roff[e % (NE / 2)] + gq];
            if constexpr (decltype(tail)::value) v = k0 + q < KI ? v : (TW)0;
            ra[e] = v;
        }
        rap[0] = xs[gq];
        rbp[0] = xs[gq + nsrc];
        ram[0] = xs[-gq];
        rbm[0] = xs[-gq + nsrc];
        rph[0] = ps[gq];
    };
    auto load_step = [&](int k0) {
        if (k0 + FK <= KI)
            load(k0, std::false_type{});
        else
            load(k0, std::true_type{});
    };
    // MFMA tiles wholly in the padding of the last row, column or inner tile are
    // skipped per wave: live 16-row tiles and 16-column tiles of this wave
    const int ni = M - row0 >= RM ? NI : (M - row0 + 15) / 16;
    const int ncw = NC - col0 - wc;
    const int nj = ncw >= 32 ? 2 : (ncw <= 0 ? 0 : (ncw + 15) / 16);
    auto step = [&](int k0, auto full) {
#pragma unroll
        for (int kc = 0; kc < FK; kc += 4) {
            if constexpr (!decltype(full)::value) {
                if (k0 + kc >= KI) break;
            }
            TC a[NI], b[2];
#pragma unroll
            for (int i = 0; i < NI; ++i) a[i] = As[kc + lk][g * RM + 16 * i + li];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[g][kc + lk][wc + 16 * j + li];
#pragma unroll
            for (int i = 0; i < NI; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if constexpr (!decltype(full)::value) {
                        if (i >= ni || j >= nj) continue;
                    }
                    acc[i][j] = Mx<TC>::mma(a[i], b[j], acc[i][j]);
                }
        }
    };
    load_step(0);
    for (int k0 = 0; k0 < KI; k0 += FK) {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            int idx = tid + 256 * e;
            As[idx & 15][idx >> 4] = (TC)ra[e];
        }
#pragma unroll
        for (int e = 0; e < 1; ++e) {
            int idx = tid + 256 * e;
            int t = idx >> 4, q = idx & 15;
            int m = k0 + q;
            double2 p = rph[e];                          // e^{i m phi}
            // u_m (+ and - parts) and u_{-m} (phase e^{-i m phi})
            double xpr = rap[e].x + rbp[e].x, xpi = rap[e].y + rbp[e].y;
            double xmr = rap[e].x - rbp[e].x, xmi = rap[e].y - rbp[e].y;
            double ypr = ram[e].x + rbm[e].x, ypi = ram[e].y + rbm[e].y;
            double ymr = ram[e].x - rbm[e].x, ymi = ram[e].y - rbm[e].y;
            double u0r = p.x * xpr - p.y * xpi, u0i = p.x * xpi + p.y * xpr;   // u+_m
            double u1r = p.x * xmr - p.y * xmi, u1i = p.x * xmi + p.y * xmr;   // u-_m
            double v0r = p.x * ypr + p.y * ypi, v0i = p.x * ypi - p.y * ypr;   // u+_{-m}
            double v1r = p.x * ymr + p.y * ymi, v1i = p.x * ymi - p.y * ymr;   // u-_{-m}
            double sg = (m & 1) ? -1.0 : 1.0;
            if (m == 0) {
                sg = 0.0;  // s_0 = u_0, t_0 = 0 (O has a zero column 0)
            }
            if (C.mirror && (m & 1)) {  // (-1)^m on s_m and t_m
                u0r = -u0r, u0i = -u0i, u1r = -u1r, u1i = -u1i;
                v0r = -v0r, v0i = -v0i, v1r = -v1r, v1i = -v1i;
            }
            Bs[0][q][4 * t + 0] = (TC)(u0r + sg * v0r);
            Bs[0][q][4 * t + 1] = (TC)(u0i + sg * v0i);
            Bs[0][q][4 * t + 2] = (TC)(u1r + sg * v1r);
            Bs[0][q][4 * t + 3] = (TC)(u1i + sg * v1i);
            Bs[1][q][4 * t + 0] = (TC)(u0r - sg * v0r);
            Bs[1][q][4 * t + 1] = (TC)(u0i - sg * v0i);
            Bs[1][q][4 * t + 2] = (TC)(u1r - sg * v1r);
            Bs[1][q][4 * t + 3] = (TC)(u1i - sg * v1i);
        }
        __syncthreads();
        if (k0 + FK < KI) load_step(k0 + FK);
        if (ni == NI && nj == 2 && k0 + FK <= KI)
            step(k0, std::true_type{});
        else if (ni > 0 && nj > 0)
            step(k0, std::false_type{});
        __syncthreads();
    }
    // epilogue: the O-product waves hand T to the E-product waves through LDS
    TC *X = smem;  // RM x (FNF2 + RPAD)
    if (g == 1) {
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) X[(16 * i + Mx<TC>::drow(v, lk)) * (FNF2 + RPAD) + wc + 16 * j + li] = acc[i][j][v];
    }
    __syncthreads();
    if (g == 1) return;
    TC *Y = wy + rows_off(nu, K) * (int64_t)NCY + 4 * (int64_t)C.ycol0 * C.nrhs;
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            int c = wc + 16 * j + li;
            int gc = col0 + c;
            if (gc >= NC) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int r = 16 * i + Mx<TC>::drow(v, lk);
                int k = row0 + r;
                if (k >= M) continue;
                TC S = acc[i][j][v], T = X[r * (FNF2 + RPAD) + c];
                if (C.mirror) {
                    TC f = ((nu + k) & 1) ? -1 : 1;
                    S *= f;
                    T *= -f;
                }
                Y[(int64_t)(kk + k) * NCY + gc] = (TC)0.5 * (S + T);
                if (k > 0) Y[(int64_t)(kk - k) * NCY + gc] = ((k & 1) ? (TC)-0.5 : (TC)0.5) * (S - T);
            }
        }
}

template <typename TC, typename TW, int FM, int MINB>
__global__ __launch_bounds__(256, MINB) void k_class_rot_back(const GroupDesc *__restrict__ desc,
                                                        const int32_t *__restrict__ pidx,
                                                        const double2 *__restrict__ phase,
                                                        const TW *__restrict__ wd, const TC *__restrict__ wz,
                                                        double2 *__restrict__ W) {
    typedef typename Mx<TC>::v4 acc4;
    const GroupDesc C = *desc;
    const int n = blockIdx.z + 1;
    if (n > C.Lt) return;
    const int K = C.K;
    const int kk = n < K ? n : K;
    const int M = n + 1;    // rows m = 0..n
    const int KI = kk + 1;  // inner k = 0..kk
    const int NTOT = C.ncols * C.nrhs;
    const int NC = 4 * NTOT;
    const int NCY = 4 * C.ncy * C.nrhs;
    const int row0 = blockIdx.y * FM;
    const int col0 = blockIdx.x * FNB;
    if (row0 >= M) return;
    __shared__ TC As[FK][2 * FM + RPAD];     // columns 0..FM-1: E^T rows, FM..2FM-1: O^T rows
    __shared__ TC Bs[2][FK][FNB + RPAD];     // s, t
    const TW *E = wd + C.off_d + fold_off(n, K);  // E[k][m], KI x M
    const TW *O = E + (int64_t)KI * M;
    const TC *Z = wz + rows_off(n, K) * (int64_t)NCY + 4 * (int64_t)C.ycol0 * C.nrhs;  // row kk + k
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    constexpr int NI = FM / 32;         // 16-row MFMA tiles of a wave (FM / 2 rows)
    constexpr int NEA = FM / 8;         // table entries a thread loads per step
    const int wr = (w >> 1) * (FM / 2);
    const int wc = (w & 1) * (FNB / 2);
    acc4 accF[NI][FNB / 32], accG[NI][FNB / 32];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < FNB / 32; ++j) {
            accF[i][j] = acc4{0, 0, 0, 0};
            accG[i][j] = acc4{0, 0, 0, 0};
        }
    // raw loads only: the s/t combination happens at the LDS store, so that the
    // prefetch of the next step stays in flight while the matrix cores work
    TW ra[NEA];
    TC rzp[FNB / 16], rzm[FNB / 16];
    // Loads without bounds branches.  A thread's table row m (tid & 63 of E or O) and
    // Z column (tid & 63) are clamped into the operands: rows m >= M and columns
    // >= NC only feed outputs that are never stored.  Inner entries k >= KI occur in
    // the last step only, which clamps them and zeroes their table entries; z_{-0}
    // is zeroed at the LDS store of the first step.
    const TW *pa = ((tid & (2 * FM - 1)) < FM ? E : O) + min(row0 + (tid & (FM - 1)), M - 1);
    const TC *pz = Z + min(col0 + (tid & (FNB - 1)), NC - 1);
    auto load = [&](int k0, auto tail) {
#pragma unroll
        for (int e = 0; e < NEA; ++e) {
            const int gk = k0 + tid / (2 * FM) + (256 / (2 * FM)) * e;  // 16 inner x 2 FM rows (E, O), coalesced along m
            if constexpr (decltype(tail)::value) {
                const TW v = pa[min(gk, KI - 1) * M];
                ra[e] = gk < KI ? v : (TW)0;
            } else {
                ra[e] = pa[gk * M];
            }
        }
#pragma unroll
        for (int e = 0; e < FNB / 16; ++e) {
            int gk = k0 + (tid >> 6) + 4 * e;  // 16 inner x FNB columns
            if constexpr (decltype(tail)::value) gk = min(gk, KI - 1);
            rzp[e] = pz[(kk + gk) * NCY];
            rzm[e] = pz[(kk - gk) * NCY];
        }
    };
    auto load_step = [&](int k0) {
        if (k0 + FK <= KI)
            load(k0, std::false_type{});
        else
            load(k0, std::true_type{});
    };
    // MFMA tiles wholly in the padding of the last row, column or inner tile are
    // skipped per wave: live 16-row tiles and 16-column tiles of this wave
    const int nrw = M - row0 - wr;
    const int ni = nrw >= FM / 2 ? NI : (nrw <= 0 ? 0 : (nrw + 15) / 16);
    const int ncw = NC - col0 - wc;
    const int nj = ncw >= FNB / 2 ? FNB / 32 : (ncw <= 0 ? 0 : (ncw + 15) / 16);
    auto step = [&](int k0, auto full) {
#pragma unroll
        for (int kc = 0; kc < FK; kc += 4) {
            if constexpr (!decltype(full)::value) {
                if (k0 + kc >= KI) break;
            }
            TC ae[NI], ao[NI], bs[FNB / 32], bt[FNB / 32];
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                ae[i] = As[kc + lk][wr + 16 * i + li];
                ao[i] = As[kc + lk][FM + wr + 16 * i + li];
            }
#pragma unroll
            for (int j = 0; j < FNB / 32; ++j) {
                bs[j] = Bs[0][kc + lk][wc + 16 * j + li];
                bt[j] = Bs[1][kc + lk][wc + 16 * j + li];
            }
#pragma unroll
            for (int i = 0; i < NI; ++i)
#pragma unroll
                for (int j = 0; j < FNB / 32; ++j) {
                    if constexpr (!decltype(full)::value) {
                        if (i >= ni || j >= nj) continue;
                    }
                    accF[i][j] = Mx<TC>::mma(ae[i], bs[j], accF[i][j]);
                    accG[i][j] = Mx<TC>::mma(ao[i], bt[j], accG[i][j]);
                }
        }
    };
    load_step(0);
    for (int k0 = 0; k0 < KI; k0 += FK) {
#pragma unroll
        for (int e = 0; e < NEA; ++e) {
            int idx = tid + 256 * e;
            As[idx / (2 * FM)][idx & (2 * FM - 1)] = (TC)ra[e];
        }
#pragma unroll
        for (int e = 0; e < FNB / 16; ++e) {
            int idx = tid + 256 * e;
            int q = idx / FNB, c = idx % FNB;
            int gk = k0 + q;
            // s_k = z_k + (-1)^k z_{-k}, t_k = z_k - (-1)^k z_{-k}; s_0 = z_0, t_0 = 0
            TC zm = rzm[e];
            if (e == 0 && k0 == 0) zm = q == 0 ? (TC)0 : zm;  // the load read z_0 for z_{-0}
            TC sg = (gk & 1) ? -1 : 1;
            TC f = (C.mirror && (gk & 1)) ? -1 : 1;  // (-1)^k on s_k and t_k
            Bs[0][q][c] = f * (rzp[e] + sg * zm);
            Bs[1][q][c] = f * (rzp[e] - sg * zm);
        }
        __syncthreads();
        if (k0 + FK < KI) load_step(k0 + FK);
        if (ni == NI && nj == FNB / 32 && k0 + FK <= KI)
            step(k0, std::true_type{});
        else if (ni > 0 && nj > 0)
            step(k0, std::false_type{});
        __syncthreads();
    }
    // w_m and w_{-m}; the 4 parts of a translation sit in lanes li..li+3 (li % 4 == 0)
    const int64_t ndst = (int64_t)C.Lt * (C.Lt + 2);
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < FNB / 32; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int c = wc + 16 * j + li;
                double F = (double)accF[i][j][v], G = (double)accG[i][j][v];
                if (C.mirror) {
                    double f = ((n + row0 + wr + 16 * i + Mx<TC>::drow(v, lk)) & 1) ? -1.0 : 1.0;  // (-1)^{n+m}
                    F *= f;
                    G *= -f;
                }
                double wp = 0.5 * (F + G), wm = 0.5 * (F - G);
                double p1 = __shfl_down(wp, 1, 64), p2 = __shfl_down(wp, 2, 64), p3 = __shfl_down(wp, 3, 64);
                double q1 = __shfl_down(wm, 1, 64), q2 = __shfl_down(wm, 2, 64), q3 = __shfl_down(wm, 3, 64);
                int gc = col0 + c;
                int m = row0 + wr + 16 * i + Mx<TC>::drow(v, lk);
                if ((li & 3) != 0 || gc >= NC || m >= M) continue;
                int t = gc >> 2;
                int tt = t % C.ncols, rr = t / C.ncols;  // translation, right-hand side
                double2 ph = phase[(int64_t)pidx[C.col_begin + tt] * (2 * C.PL + 1) + C.PL + m];  // e^{i m phi}
                double2 *dst = W + ((int64_t)rr * C.ncy + C.ycol0 + tt) * (2 * ndst);
                {
                    // w_m: parts (wp, p1) = z_+, (p2, p3) = z_-; a = (z_+ + z_-)/2, b = (z_+ - z_-)/2; phase e^{-i m phi}
                    double ar = 0.5 * (wp + p2), ai = 0.5 * (p1 + p3);
                    double br = 0.5 * (wp - p2), bi = 0.5 * (p1 - p3);
                    int64_t o = fidx(n, m);
                    dst[o] = make_double2(ph.x * ar + ph.y * ai, ph.x * ai - ph.y * ar);
                    dst[o + ndst] = make_double2(ph.x * br + ph.y * bi, ph.x * bi - ph.y * br);
                }
                if (m > 0) {
                    // w_{-m} = (-1)^m (F - G)/2; phase e^{+i m phi}
                    double sg = (m & 1) ? -1.0 : 1.0;
                    double ar = 0.5 * sg * (wm + q2), ai = 0.5 * sg * (q1 + q3);
                    double br = 0.5 * sg * (wm - q2), bi = 0.5 * sg * (q1 - q3);
                    int64_t o = fidx(n, -m);
                    dst[o] = make_double2(ph.x * ar - ph.y * ai, ph.x * ai + ph.y * ar);
                    dst[o + ndst] = make_double2(ph.x * br - ph.y * bi, ph.x * bi + ph.y * br);
                }
            }
}

// Coaxial shift of one segment: grid (translation tiles, row tiles, 2K+1 orders).
//   Z^+ = P Y^+,  Z^- = Q Y^-  (complex), computed as real products of
//   [[Pr, -Pi], [Pi, Pr]] with [Yr; Yi].  One 16x16x4 tile = 8 rows n x 2 parts
//   (tile row r: n = r % 8, part = r / 8), inner = 2 nu x 2 parts, 16 translations.
//   Block: 64 rows n x 64 translations; wave w owns rows 16 w .. 16 w + 15.
constexpr int AM = 64;
constexpr int AK = 16;

// AN: translations per block (64; 32 for the last tile of a segment whose remainder
// is at most 32, which would otherwise leave half of a 64-wide tile empty); col_base:
// first translation of the launch.
template <typename TC, int AN>
__global__ __launch_bounds__(256) void k_class_axial(const GroupDesc *__restrict__ desc,
                                                     const double2 *__restrict__ wax, const TC *__restrict__ wy,
                                                     TC *__restrict__ wz, int col_base) {
    typedef typename Mx<TC>::v4 acc4;
    typedef typename V2<TC>::t c2;
    constexpr int NJ = AN / 16;        // 16-translation MFMA tiles of a wave
    constexpr int NY = AN / 4;         // Y entries a thread loads per step (16 inner x 4 AN reals)
    const GroupDesc C = *desc;
    const int K = C.K;
    const int k = (int)blockIdx.z - K;
    const int ak = k < 0 ? -k : k;
    const int q = ak < 1 ? 1 : ak;
    const int Lt = C.Lt, Ls = C.Ls;
    const int M = Lt - q + 1;
    const int KI = Ls - q + 1;
    const int row0 = blockIdx.y * AM;
    const int col0 = col_base + blockIdx.x * AN;
    if (row0 >= M) return;
    const int NC = 4 * C.ncy * C.nrhs;                     // row stride of Y and Z
    const int64_t cb = 4 * (int64_t)C.ycol0 * C.nrhs;      // first real column of the segment
    const int NTOT = C.ncols * C.nrhs;                     // columns of the segment
    const double2 *Pb = wax + C.off_ax + ax_off(k, Lt, Ls, K);
    const double2 *Qb = Pb + ax_block(k, Lt, Ls);
    // rows padded by one entry: the tile store writes 8 inner columns of one row from
    // 8 consecutive lanes, which a 1 KiB row stride would put on the same banks
    __shared__ c2 Ps[AK][AM + 1];
    __shared__ c2 Qs[AK][AM + 1];
    // Y tile by part: the 16 lanes of an MFMA B fragment read 16 consecutive columns of
    // one part.  Part p of an inner row starts YP(p) = 80 p (+ 8 for p >= 2) doubles in,
    // i.e. at bank offsets 0, 32, 16, 48 of the 64 four-byte banks: the two parts one
    // half-wave reads (0,1 or 2,3) lie 32 banks apart, and the tile store (8 columns x
    // 4 parts per half-wave) covers all 64 banks once; with parts 80 doubles apart,
    // parts 0 and 2 shared banks and the store was a 2-way conflict.
    constexpr int YROW = 3 * 80 + 8 + AN;
    __shared__ TC Ys[AK][YROW];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    // A-fragment row li -> (n within 8, output part), chosen so that values h and h + 2 of
    // a lane are the (re, im) parts of one n in both D layouts: FP64 row 4v + lk -> n = 4h + lk;
    // FP32 row 4 lk + v -> n = h + 2 lk
    const int nl = Mx<TC>::F32 ? ((li & 1) | ((li >> 2) << 1)) : (li & 7);
    const int pout = Mx<TC>::F32 ? ((li >> 1) & 1) : (li >> 3);
    const int nul = lk >> 1, pin = lk & 1;   // inner: (nu within 2, input part)
    const bool use_im = pout != pin;
    const TC sgn = (pout == 0 && pin == 1) ? -1 : 1;
    acc4 accP[2][NJ], accQ[2][NJ];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            accP[i][j] = acc4{0, 0, 0, 0};
            accQ[i][j] = acc4{0, 0, 0, 0};
        }
    double2 pa[4], pb[4];
    TC py[NY];
    const int ycol = AN == 64 ? tid : tid % (4 * AN);   // this thread's real column of the tile
    const int yrow = AN == 64 ? 0 : tid / (4 * AN);     // ... and its first inner row
    const bool ycol_ok = col0 + (ycol >> 2) < NTOT;      // its translation column
    const TC *ybase = wy + cb + 4 * (int64_t)col0 + ycol;
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;  // 64 rows x 16 inner
            int r = idx >> 4, c = idx & 15;
            int gr = row0 + r, gc = k0 + c;
            bool ok = gr < M && gc < KI;
            pa[e] = ok ? Pb[(int64_t)gr * KI + gc] : make_double2(0.0, 0.0);
            pb[e] = ok ? Qb[(int64_t)gr * KI + gc] : make_double2(0.0, 0.0);
        }
        // inner rows nu = q + k0 + yrow + (256 / (4 AN)) e x 4 AN reals (AN translations x 4);
        // the row of (nu, k) advances by 2 nu + 2 (nu < K) or 2 K + 1 per degree
        int nu = q + k0 + yrow;
        int64_t row = rows_off(nu, K) + k + (nu < K ? nu : K);
#pragma unroll
        for (int e = 0; e < NY; ++e) {
            py[e] = (ycol_ok && nu <= Ls) ? ybase[row * NC] : (TC)0;
            row += nu < K ? 2 * nu + 2 : 2 * K + 1;
            ++nu;
            if constexpr (AN < 64) {  // inner rows 256 / (4 AN) apart
#pragma unroll
                for (int st = 1; st < 256 / (4 * AN); ++st) {
                    row += nu < K ? 2 * nu + 2 : 2 * K + 1;
                    ++nu;
                }
            }
        }
    };
    // A wave whose 16 rows lie wholly in the padding of the last row tile skips the
    // products (a uniform branch), which leaves the matrix core to the other resident waves
    const bool live = M - row0 - 16 * w > 0;
    auto step = [&]() {
#pragma unroll
        for (int cc = 0; cc < AK; cc += 2) {
            TC aP[2], aQ[2], bP[NJ], bQ[NJ];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                int r = w * 16 + 8 * i + nl;
                c2 p = Ps[cc + nul][r], qv = Qs[cc + nul][r];
                aP[i] = sgn * (use_im ? p.y : p.x);
                aQ[i] = sgn * (use_im ? qv.y : qv.x);
            }
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                int col = 16 * j + li;
                bP[j] = Ys[cc + nul][80 * pin + col];
                bQ[j] = Ys[cc + nul][168 + 80 * pin + col];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    accP[i][j] = Mx<TC>::mma(aP[i], bP[j], accP[i][j]);
                    accQ[i][j] = Mx<TC>::mma(aQ[i], bQ[j], accQ[i][j]);
                }
        }
    };
    load(0);
    for (int k0 = 0; k0 < KI; k0 += AK) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int idx = tid + 256 * e;
            Ps[idx & 15][idx >> 4] = c2{(TC)pa[e].x, (TC)pa[e].y};
            Qs[idx & 15][idx >> 4] = c2{(TC)pb[e].x, (TC)pb[e].y};
        }
#pragma unroll
        for (int e = 0; e < NY; ++e) {
            if constexpr (AN == 64)
                Ys[e][80 * (tid & 3) + ((tid & 2) ? 8 : 0) + (tid >> 2)] = py[e];
            else
                Ys[yrow + (256 / (4 * AN)) * e][80 * (ycol & 3) + ((ycol & 2) ? 8 : 0) + (ycol >> 2)] = py[e];
        }
        __syncthreads();
        if (k0 + AK < KI) load(k0 + AK);
        if (live) step();
        __syncthreads();
    }
    // values h and h + 2 of a lane (h = 0, 1) are the (re, im) parts of one n of tile i
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int n = q + row0 + w * 16 + 8 * i + (Mx<TC>::F32 ? h + 2 * lk : 4 * h + lk);
            if (n > Lt) continue;
            int kq = n < K ? n : K;
            TC *row = wz + (rows_off(n, K) + (k + kq)) * (int64_t)NC + cb;
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                int t = col0 + 16 * j + li;
                if (t >= NTOT) continue;
                row[4 * t + 0] = accP[i][j][h];
                row[4 * t + 1] = accP[i][j][h + 2];
                row[4 * t + 2] = accQ[i][j][h];
                row[4 * t + 3] = accQ[i][j][h + 2];
            }
        }
}

// Deterministic per-target sum of back-rotated columns: grid (targets, element blocks).
// grid (targets, element blocks, right-hand sides).
__global__ __launch_bounds__(256) void k_scatter(const ScatterSeg *__restrict__ segs, const int32_t *__restrict__ cols,
                                                 const double2 *__restrict__ W, int64_t nbox,
                                                 double2 *__restrict__ dst_pool, int ncy, int64_t dst_stride) {
    const ScatterSeg S = segs[blockIdx.x];
    const int r = blockIdx.z;
    double2 *dst = dst_pool + r * dst_stride + S.dst;
    const double2 *Wr = W + (int64_t)r * ncy * nbox;
    for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < nbox; e += (int64_t)blockDim.x * gridDim.y) {
        double2 acc = dst[e];
        for (int i = S.beg; i < S.end; ++i) {
            double2 v = Wr[(int64_t)cols[i] * nbox + e];
            acc.x += v.x;
            acc.y += v.y;
        }
        dst[e] = acc;
    }
}

template <typename TC, typename TW>
static void launch_class_t(const GroupDesc &h, const GroupDesc *d_desc, const int64_t *src_off, const int32_t *pidx,
                           const double2 *phase, const double2 *src_pool, const TW *wd, const double2 *wax,
                           TC *wy, TC *wz, int stage, int rows, hipStream_t s) {
    const int NC = 4 * h.ncols * h.nrhs;
    const int K = h.K;
    if (stage == 0) {
        const int rm = rows == 32 ? 32 : FMF;
        int mt = ((h.Ls < K ? h.Ls : K) + 1 + rm - 1) / rm;
        dim3 grid((NC + FNF2 - 1) / FNF2, mt, h.Ls);
        if (rm == 32)
            hipLaunchKernelGGL((k_class_rot_fwd<TC, TW, 32, 3>), grid, dim3(256), 0, s, d_desc, src_off, pidx, phase,
                               src_pool, wd, wy);
        else
            hipLaunchKernelGGL((k_class_rot_fwd<TC, TW, FMF, 3>), grid, dim3(256), 0, s, d_desc, src_off, pidx, phase,
                               src_pool, wd, wy);
    } else if (stage == 1) {
        int mt = (h.Lt + AM - 1) / AM;
        const int nt = h.ncols * h.nrhs;
        int nfull;
        bool tail;
        axial_tile_split(nt, nfull, tail);
        if (nfull > 0)
            hipLaunchKernelGGL((k_class_axial<TC, 64>), dim3(nfull, mt, 2 * K + 1), dim3(256), 0, s, d_desc, wax,
                               (const TC *)wy, wz, 0);
        if (tail)  // a remainder of at most 32 translations
            hipLaunchKernelGGL((k_class_axial<TC, 32>), dim3(1, mt, 2 * K + 1), dim3(256), 0, s, d_desc, wax,
                               (const TC *)wy, wz, nfull * 64);
    } else {
        const int fm = rows == 32 ? 32 : FMB;
        int mt = (h.Lt + 1 + fm - 1) / fm;
        dim3 grid((NC + FNB - 1) / FNB, mt, h.Lt);
        // W (FP64 columns) aliases the Y workspace, free once the coaxial shifts are done
        if (fm == 32)
            hipLaunchKernelGGL((k_class_rot_back<TC, TW, 32, 3>), grid, dim3(256), 0, s, d_desc, pidx, phase, wd,
                               (const TC *)wz, (double2 *)wy);
        else
            hipLaunchKernelGGL((k_class_rot_back<TC, TW, FMB, 3>), grid, dim3(256), 0, s, d_desc, pidx, phase, wd,
                               (const TC *)wz, (double2 *)wy);
    }
}

void launch_class(const GroupDesc &h, const GroupDesc *d_desc, const int64_t *src_off, const int32_t *pidx,
                  const double2 *phase, const double2 *src_pool, const double *wd, const double2 *wax, double *wy,
                  double *wz, int stage, int prec, int ftab, int rows, hipStream_t s) {
    if (h.ncols <= 0 || h.nrhs <= 0) return;
    if (prec == 1 && ftab)
        launch_class_t<float, float>(h, d_desc, src_off, pidx, phase, src_pool, (const float *)wd, wax, (float *)wy,
                                     (float *)wz, stage, rows, s);
    else if (prec == 1)
        launch_class_t<float, double>(h, d_desc, src_off, pidx, phase, src_pool, wd, wax, (float *)wy, (float *)wz,
                                      stage, rows, s);
    else
        launch_class_t<double, double>(h, d_desc, src_off, pidx, phase, src_pool, wd, wax, wy, wz, stage, rows, s);
}

void launch_scatter(const ScatterSeg *segs, int nsegs, const int32_t *cols, const double2 *W, int64_t nbox,
                    double2 *dst_pool, int ncy, int nrhs, int64_t dst_stride, hipStream_t s) {
    if (nsegs <= 0 || nrhs <= 0) return;
    int yb = (int)std::min<int64_t>((nbox + 255) / 256, 64);
    hipLaunchKernelGGL(k_scatter, dim3(nsegs, yb, nrhs), dim3(256), 0, s, segs, cols, W, nbox, dst_pool, ncy,
                       dst_stride);
}

}  // namespace ifs
