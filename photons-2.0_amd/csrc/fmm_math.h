// fmm_math.h -- Cartesian multipole algebra, pair kernels and acceptance rules
// for the short-range (erfc-split) FMM, as device functions for gfx950.
//
// Multi-index layout: terms ordered by total degree n, then a (x power)
// descending, then b descending -- the layout of inc/operator.h:24-164, extended
// to any order.  index(a,b,c) = n(n+1)(n+2)/6 + m(m+1)/2 + c with m = b + c.
//
// Conventions (src/operator.c):
//   P2M  M_a  = sum_i m (-(x_i - c))^a / a!
//   M2M  M'_a = sum_{b<=a} M_b d^(a-b)/(a-b)!            d = c_parent - c_child
//   M2L  L_a += sum_{|a|+|b|<=p} M_b D_{a+b}(R)          R = c_target - c_source
//   L2L  L'_a += sum_g L_{a+g} d^g/g!                     d = c_child - c_parent
//   L2P  acc_i += sum_{|a|<=p-1} L_{a+e_i} d^a/a!         d = x - c_leaf
//   D_a(R) = d^a g(|R|),  g(r) = erfc(r/2rs)/r,
//   D_a = sum_k f_{|a|-k} sum_{kx+ky+kz=k} prod_i a_i!/(2^k_i k_i! (a_i-2k_i)!) R_i^(a_i-2k_i)
//   with f_n = (r^-1 d/dr)^n g.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

namespace photons {

__host__ __device__ constexpr int nterms(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
__host__ __device__ constexpr int mindex(int a, int b, int c)
{
    return (a + b + c) * (a + b + c + 1) * (a + b + c + 2) / 6 + (b + c) * (b + c + 1) / 2 + c;
}

constexpr int kMaxOrder = 5;
constexpr int kMaxTerms = nterms(kMaxOrder);

struct MultiIndexTable {
    int a[nterms(2 * kMaxOrder)], b[nterms(2 * kMaxOrder)], c[nterms(2 * kMaxOrder)];
    double inv_fact[nterms(2 * kMaxOrder)];  // 1/(a! b! c!)
    constexpr MultiIndexTable() : a(), b(), c(), inv_fact()
    {
        int k = 0;
        for (int n = 0; n <= 2 * kMaxOrder; n++)
            for (int x = n; x >= 0; x--)
                for (int y = n - x; y >= 0; y--) {
                    a[k] = x;
                    b[k] = y;
                    c[k] = n - x - y;
                    double f = 1.0;
                    for (int i = 2; i <= x; i++) f *= i;
                    for (int i = 2; i <= y; i++) f *= i;
                    for (int i = 2; i <= n - x - y; i++) f *= i;
                    inv_fact[k] = 1.0 / f;
                    k++;
                }
    }
};
constexpr MultiIndexTable kMI{};

__host__ __device__ constexpr int ma(int k) { return kMI.a[k]; }
__host__ __device__ constexpr int mb(int k) { return kMI.b[k]; }
__host__ __device__ constexpr int mc(int k) { return kMI.c[k]; }
__host__ __device__ constexpr int mdeg(int k) { return kMI.a[k] + kMI.b[k] + kMI.c[k]; }

// a!/(2^k k! (a-2k)!): coefficient of the k-th contraction of d^a applied to a radial function
__host__ __device__ constexpr double hcoef(int a, int k)
{
    double num = 1.0, den = 1.0;
    for (int i = 2; i <= a; i++) num *= i;
    for (int i = 2; i <= k; i++) den *= i;
    for (int i = 2; i <= a - 2 * k; i++) den *= i;
    for (int i = 0; i < k; i++) den *= 2.0;
    return num / den;
}

// Compile-time loop: f(std::integral_constant<int, I>) for I = B .. E-1, so that every
// multi-index below is a constant expression and the term arrays stay in registers.
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Powers x^e / e! for e = 0..P, i.e. the monomial weights of a Taylor shift.
template <int P, typename T>
__device__ __forceinline__ void scaled_powers(T x, T *px)
{
    px[0] = T(1);
    static_for<1, P + 1>([&](auto e) { px[e.value] = px[e.value - 1] * x * T(1.0 / e.value); });
}

template <int P, typename T>
__device__ __forceinline__ void plain_powers(T x, T *px)
{
    px[0] = T(1);
    static_for<1, P + 1>([&](auto e) { px[e.value] = px[e.value - 1] * x; });
}

// ----------------------------------------------------------------------------- P2M / M2M

// M += m * (-d)^a / a!  for one particle at offset d = x - c
template <int P, typename T>
__device__ __forceinline__ void p2m_add(T m, T dx, T dy, T dz, T *M)
{
    T px[P + 1], py[P + 1], pz[P + 1];
    scaled_powers<P>(-dx, px);
    scaled_powers<P>(-dy, py);
    scaled_powers<P>(-dz, pz);
    static_for<0, nterms(P)>([&](auto k) {
        constexpr int K = k.value;
        M[K] += m * px[ma(K)] * py[mb(K)] * pz[mc(K)];
    });
}

// tM += shift(M), d = c_parent - c_child
template <int P, typename T>
__device__ __forceinline__ void m2m_add(T dx, T dy, T dz, const T *M, T *tM)
{
    T px[P + 1], py[P + 1], pz[P + 1];
    scaled_powers<P>(dx, px);
    scaled_powers<P>(dy, py);
    scaled_powers<P>(dz, pz);
    static_for<0, nterms(P)>([&](auto k) {
        constexpr int K = k.value;
        T s = M[K];
        static_for<0, K>([&](auto j) {
            constexpr int J = j.value;
            if constexpr (ma(J) <= ma(K) && mb(J) <= mb(K) && mc(J) <= mc(K))
                s += M[J] * px[ma(K) - ma(J)] * py[mb(K) - mb(J)] * pz[mc(K) - mc(J)];
        });
        tM[K] += s;
    });
}

// ----------------------------------------------------------------------------- L2L / L2P

// tL += shift(L), d = c_child - c_parent
template <int P, typename T>
__device__ __forceinline__ void l2l_add(T dx, T dy, T dz, const T *L, T *tL)
{
    T px[P + 1], py[P + 1], pz[P + 1];
    scaled_powers<P>(dx, px);
    scaled_powers<P>(dy, py);
    scaled_powers<P>(dz, pz);
    static_for<0, nterms(P)>([&](auto k) {
        constexpr int K = k.value;
        T s = L[K];
        static_for<1, nterms(P - mdeg(K))>([&](auto g) {
            constexpr int G = g.value;
            s += L[mindex(ma(K) + ma(G), mb(K) + mb(G), mc(K) + mc(G))] * px[ma(G)] * py[mb(G)] * pz[mc(G)];
        });
        tL[K] += s;
    });
}

// gradient of the local expansion at offset d from its centre
template <int P, typename T>
__device__ __forceinline__ void l2p_eval(const T *L, T dx, T dy, T dz, T &ax, T &ay, T &az)
{
    T px[P + 1], py[P + 1], pz[P + 1];
    scaled_powers<P>(dx, px);
    scaled_powers<P>(dy, py);
    scaled_powers<P>(dz, pz);
    T fx = L[1], fy = L[2], fz = L[3];
    static_for<1, nterms(P - 1)>([&](auto k) {
        constexpr int K = k.value;
        T w = px[ma(K)] * py[mb(K)] * pz[mc(K)];
        fx += L[mindex(ma(K) + 1, mb(K), mc(K))] * w;
        fy += L[mindex(ma(K), mb(K) + 1, mc(K))] * w;
        fz += L[mindex(ma(K), mb(K), mc(K) + 1)] * w;
    });
    ax = fx;
    ay = fy;
    az = fz;
}

// ----------------------------------------------------------------------------- split kernel

// Single precision erfc(u) for u >= 0 as exp(-u^2) * t * Q(t), t = 1/(1 + u/2): the
// variable of the Numerical Recipes erfcc form (Press et al., sec. 6.2), with Q a
// degree-10 polynomial fitted here by least squares on the relative error of
// erfcx(u) over [0, 1e4].  Relative error of t*Q(t) against erfcx: 1.1e-8 in exact
// arithmetic, 3.2e-7 when evaluated in float (rounding limited).
__device__ __forceinline__ float erfcx_tq(float t)
{
    float q = 4.242083445e-02f;
    q = __builtin_fmaf(q, t, -2.311667988e-01f);
    q = __builtin_fmaf(q, t, 4.867675762e-01f);
    q = __builtin_fmaf(q, t, -4.475895600e-01f);
    q = __builtin_fmaf(q, t, 1.274673754e-01f);
    q = __builtin_fmaf(q, t, -5.737388979e-02f);
    q = __builtin_fmaf(q, t, 9.303219725e-02f);
    q = __builtin_fmaf(q, t, 1.753727219e-01f);
    q = __builtin_fmaf(q, t, 2.468809031e-01f);
    q = __builtin_fmaf(q, t, 2.820938369e-01f);
    q = __builtin_fmaf(q, t, 2.820947951e-01f);
    return t * q;
}

// exp(-u^2) in float via the hardware base-2 exponential
__device__ __forceinline__ float gauss_f32(float u)
{
    return __builtin_amdgcn_exp2f(-1.4426950408889634f * u * u);
}

// erfcx(u) = t Q20(t), t = 1/(1+u/2), u >= 0: a least-squares fit of the relative
// error done here in 50-digit arithmetic on 1200 Chebyshev nodes in t (relative error
// 2.6e-14 evaluated in double over u >= 0)
__device__ __forceinline__ double erfcx_tq20(double t)
{
    constexpr double c[21] = {
        2.82094791773874920e-01,  2.82094791776525355e-01,  2.46832942452028026e-01,  1.76309262185868731e-01,
        8.37464818453457355e-02,  -4.40302425728239695e-03, -5.67611432847028585e-02, -5.38614146130594953e-02,
        3.03490794480198913e-03,  -7.70416507730420971e-03, 3.74060151838098820e-01,  -1.28536904212529279e+00,
        3.66391328683279927e+00,  -8.19524649206594979e+00, 1.29420464606019987e+01,  -1.42835478387177517e+01,
        1.10704731945007335e+01,  -5.95597239466930795e+00, 2.13201704144582260e+00,  -4.58832900301043400e-01,
        4.50751019138224407e-02};
    double q = c[20];
    static_for<0, 20>([&](auto k) { q = fma(q, t, c[19 - k.value]); });
    return q;
}

// G(u) = erfc(u) + 2u/sqrt(pi) exp(-u^2) = exp(-u^2) (erfcx(u) + 2u/sqrt(pi)): one exp
// instead of the two inside erfc and beside it; t = 1/(1+u/2) by v_rcp_f64 and two
// Newton steps.  Relative error ~3e-14 against the erfc/exp form.  The FP64 P2P kernel
// takes it beyond the end of its table (u >= 6), where the table's G is zero but the
// reference's is not.
__device__ __forceinline__ double split_factor_f64(double u)
{
    const double x = fma(0.5, u, 1.0);
    double t = __builtin_amdgcn_rcp(x);
    t = fma(t, fma(-x, t, 1.0), t);
    t = fma(t, fma(-x, t, 1.0), t);
    return exp(-u * u) * fma(1.1283791670955126, u, t * erfcx_tq20(t));
}

// radial factors f_n = (r^-1 d/dr)^n [erfc(r/2rs)/r], n = 0..P, by the upward
// recurrence B_n = ((2n-1) B_{n-1} + (2a^2)^n exp(-a^2 r^2)/(a sqrt(pi))) / r^2,
// f_n = (-1)^n B_n, a = 1/(2 rs).
template <int P>
__device__ __forceinline__ void radial_f64(double r, double inv2rs, double *f)
{
    double u = r * inv2rs;
    double e = exp(-u * u) * 0.56418958354775628695;  // 1/sqrt(pi)
    double ir = 1.0 / r;
    double ir2 = ir * ir;
    double b = erfc(u) * ir;
    double t = e / inv2rs;
    double a2 = 2.0 * inv2rs * inv2rs;
    f[0] = b;
    static_for<1, P + 1>([&](auto nn) {
        constexpr int n = nn.value;
        t *= a2;
        b = ((2 * n - 1) * b + t) * ir2;
        f[n] = (n & 1) ? -b : b;
    });
}

// same in float, in units where rs = 1 (so a = 1/2)
template <int P>
__device__ __forceinline__ void radial_f32_unit(float r, float *f)
{
    float u = 0.5f * r;
    float e = gauss_f32(u);
    float ir = __builtin_amdgcn_rcpf(r);
    float ir2 = ir * ir;
    float tt = __builtin_amdgcn_rcpf(__builtin_fmaf(0.5f, u, 1.0f));
    float b = e * erfcx_tq(tt) * ir;
    float t = e * (2.0f * 0.56418958354775628695f);
    f[0] = b;
    static_for<1, P + 1>([&](auto nn) {
        constexpr int n = nn.value;
        t *= 0.5f;
        b = ((2 * n - 1) * b + t) * ir2;
        f[n] = (n & 1) ? -b : b;
    });
}

// D_a(R) for |a| <= P from the radial factors
template <int P, typename T>
__device__ __forceinline__ void dtensor(T x, T y, T z, const T *f, T *D)
{
    T px[P + 1], py[P + 1], pz[P + 1];
    plain_powers<P>(x, px);
    plain_powers<P>(y, py);
    plain_powers<P>(z, pz);
    static_for<0, nterms(P)>([&](auto k) {
        constexpr int K = k.value;
        constexpr int A = ma(K), B = mb(K), C = mc(K);
        T s = T(0);
        static_for<0, A / 2 + 1>([&](auto kx) {
            static_for<0, B / 2 + 1>([&](auto ky) {
                static_for<0, C / 2 + 1>([&](auto kz) {
                    constexpr int X = kx.value, Y = ky.value, Z = kz.value;
                    constexpr double coef = hcoef(A, X) * hcoef(B, Y) * hcoef(C, Z);
                    s += f[A + B + C - X - Y - Z] * T(coef) * px[A - 2 * X] * py[B - 2 * Y] * pz[C - 2 * Z];
                });
            });
        });
        D[K] = s;
    });
}

// L += sum_{|k|+|j|<=P} M_j D_{k+j}
template <int P, typename T>
__device__ __forceinline__ void m2l_contract(const T *D, const T *M, T *L)
{
    static_for<0, nterms(P)>([&](auto k) {
        constexpr int K = k.value;
        T s = L[K];
        static_for<0, nterms(P - mdeg(K))>([&](auto j) {
            constexpr int J = j.value;
            s += M[J] * D[mindex(ma(K) + ma(J), mb(K) + mb(J), mc(K) + mc(J))];
        });
        L[K] = s;
    });
}

// ----------------------------------------------------------------------------- acceptance

// acceptance() of src/fmm.c:267-326 (LONGSHORT): 1 accept (M2L), 0 open, -1 beyond
// the cutoff.  Boxes whose gap sums to less than 1e-4 (kpc/h) count as touching.
__host__ __device__ inline int mac(const double *wi, const double *wj, double d0, double d1, double d2,
                                   double theta2, double rc2)
{
    double w0 = (wi[0] + wj[0]) * 0.5, w1 = (wi[1] + wj[1]) * 0.5, w2 = (wi[2] + wj[2]) * 0.5;
    double dd2 = d0 * d0 + d1 * d1 + d2 * d2;
    double g0 = (d0 < 0.0 ? -d0 : d0) - w0;
    double g1 = (d1 < 0.0 ? -d1 : d1) - w1;
    double g2 = (d2 < 0.0 ? -d2 : d2) - w2;
    g0 = g0 <= 0.0 ? 0.0 : g0;
    g1 = g1 <= 0.0 ? 0.0 : g1;
    g2 = g2 <= 0.0 ? 0.0 : g2;
    if (g0 + g1 + g2 < 0.0001) return 0;
    double dm2 = g0 * g0 + g1 * g1 + g2 * g2;
    if (dm2 >= rc2) return -1;
    if (dd2 > rc2) return 0;
    double wmax = w0;
    if (w1 > wmax) wmax = w1;
    if (w2 > wmax) wmax = w2;
    wmax *= 2;
    return (wmax * wmax < theta2 * dd2) ? 1 : 0;
}

}  // namespace photons
