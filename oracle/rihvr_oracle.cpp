// rihvr_oracle.cpp -- CPU restatement of the reference's sparse regularized
// inverse reconstruction (FASTA/FISTA + fused lasso).
//
// TEST INFRASTRUCTURE ONLY: this file is the parity oracle and the CPU baseline.
// The product library never links, loads or calls it.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off -fopenmp).  -ffp-contract=off
// keeps every a*b+c as two roundings so the float evaluation order written here is
// the one executed.
//
// Each function cites the reference file:line it restates.  Template parameter T:
//   float  -> the reference's FP32 semantics (see rihvr_oracle.h)
//   double -> the FP64 oracle (no float rounding of phase arguments, double sums)
#include "rihvr_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;  // M_PI

template <class T> struct C2 { T x, y; };

// ---------------------------------------------------------------------------
// Mixed-radix Stockham FFT (own implementation; stands in for cuFFT, whose
// arithmetic is not in the reference tree).  Unnormalised, sign -1 = forward.
// Radices: 4, 2, then every odd prime factor (generic O(p^2) DFT), so any
// length is supported.  Twiddles are computed in double and rounded to T.
// `s` interleaved transforms are processed at once (element q of transform
// ... is at q + s*j), which makes the column pass of a 2D FFT a single call.
// ---------------------------------------------------------------------------
template <class T> struct FFT1 {
    int n = 0;
    std::vector<int> rad;
    std::vector<std::vector<C2<T>>> tw[2];    // [sign<0 ? 0 : 1][stage][j], exp(sign*2*pi*i*j/n_stage)
    std::vector<std::vector<C2<T>>> root[2];  // [..][stage][j], exp(sign*2*pi*i*j/p)

    explicit FFT1(int n_) : n(n_) {
        int m = n;
        while (m % 4 == 0) { rad.push_back(4); m /= 4; }
        while (m % 2 == 0) { rad.push_back(2); m /= 2; }
        for (int p = 3; m > 1; p += 2) {
            while (m % p == 0) { rad.push_back(p); m /= p; }
            if (p * p > m && m > 1) { rad.push_back(m); m = 1; }
        }
        for (int si = 0; si < 2; ++si) {
            double sign = si == 0 ? -1.0 : 1.0;
            int ncur = n;
            for (int p : rad) {
                std::vector<C2<T>> t(ncur), r(p);
                for (int j = 0; j < ncur; ++j) {
                    double a = sign * 2.0 * kPi * (double)j / (double)ncur;
                    t[j] = {(T)std::cos(a), (T)std::sin(a)};
                }
                for (int j = 0; j < p; ++j) {
                    double a = sign * 2.0 * kPi * (double)j / (double)p;
                    r[j] = {(T)std::cos(a), (T)std::sin(a)};
                }
                tw[si].push_back(std::move(t));
                root[si].push_back(std::move(r));
                ncur /= p;
            }
        }
    }

    static inline C2<T> cmul(C2<T> a, C2<T> b) {
        return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
    }

    // x: n*s elements (in/out), work: n*s scratch
    void exec(C2<T>* x, C2<T>* work, int s, int sign) const {
        const int si = sign < 0 ? 0 : 1;
        C2<T>* a = x;
        C2<T>* b = work;
        int ncur = n, stride = s;
        C2<T> in[64], out[64];
        std::vector<C2<T>> inb, outb;
        for (size_t st = 0; st < rad.size(); ++st) {
            const int p = rad[st];
            const int m = ncur / p;
            const C2<T>* tw_ = tw[si][st].data();
            const C2<T>* rt = root[si][st].data();
            C2<T>* pin = in;
            C2<T>* pout = out;
            if (p > 64) { inb.resize(p); outb.resize(p); pin = inb.data(); pout = outb.data(); }
            for (int i = 0; i < m; ++i) {
                for (int q = 0; q < stride; ++q) {
                    for (int r = 0; r < p; ++r) pin[r] = a[q + (size_t)stride * (i + r * m)];
                    dft(p, pin, pout, rt, sign);
                    for (int k = 0; k < p; ++k) {
                        C2<T> v = pout[k];
                        if (i * k != 0) v = cmul(v, tw_[i * k]);
                        b[q + (size_t)stride * (p * i + k)] = v;
                    }
                }
            }
            ncur = m;
            stride *= p;
            std::swap(a, b);
        }
        if (a != x) std::memcpy(x, a, sizeof(C2<T>) * (size_t)n * s);
    }

    static inline void dft(int p, const C2<T>* a, C2<T>* c, const C2<T>* rt, int sign) {
        if (p == 2) {
            c[0] = {a[0].x + a[1].x, a[0].y + a[1].y};
            c[1] = {a[0].x - a[1].x, a[0].y - a[1].y};
            return;
        }
        if (p == 4) {
            C2<T> t0 = {a[0].x + a[2].x, a[0].y + a[2].y};
            C2<T> t1 = {a[0].x - a[2].x, a[0].y - a[2].y};
            C2<T> t2 = {a[1].x + a[3].x, a[1].y + a[3].y};
            C2<T> t3 = {a[1].x - a[3].x, a[1].y - a[3].y};
            // w4 * t3 with w4 = -i (forward) or +i (inverse)
            C2<T> u = sign < 0 ? C2<T>{t3.y, -t3.x} : C2<T>{-t3.y, t3.x};
            c[0] = {t0.x + t2.x, t0.y + t2.y};
            c[2] = {t0.x - t2.x, t0.y - t2.y};
            c[1] = {t1.x + u.x, t1.y + u.y};
            c[3] = {t1.x - u.x, t1.y - u.y};
            return;
        }
        for (int k = 0; k < p; ++k) {
            C2<T> acc = a[0];
            for (int r = 1; r < p; ++r) {
                C2<T> w = rt[(r * k) % p];
                acc.x += a[r].x * w.x - a[r].y * w.y;
                acc.y += a[r].x * w.y + a[r].y * w.x;
            }
            c[k] = acc;
        }
    }
};

template <class T> struct FFT2 {
    int H, W;
    FFT1<T> fw, fh;
    FFT2(int H_, int W_) : H(H_), W(W_), fw(W_), fh(H_) {}
    // work: H*W elements
    void exec(C2<T>* d, C2<T>* work, int sign) const {
        for (int y = 0; y < H; ++y) fw.exec(d + (size_t)y * W, work, 1, sign);
        fh.exec(d, work, W, sign);
    }
};

// ---------------------------------------------------------------------------
// Parameters and tables
// ---------------------------------------------------------------------------
// plane_list[i] = start_plane + i * plane_stepsize, in float
// (src/core/holo_ui.cpp:478-483).
std::vector<float> plane_list(const rihvr_params* p) {
    std::vector<float> z(p->num_planes);
    for (int i = 0; i < p->num_planes; ++i) {
        float step = (float)i * p->plane_stepsize;
        z[i] = p->start_plane + step;
    }
    return z;
}

// kappa(x, y) = -2*pi*sqrt(max(0, 1 - f^2 (lambda/dx)^2)) / lambda, unshifted
// FFT index order (sparse_compressive_holo.cu:38-56; sparse_gradient.cu:27-45).
std::vector<double> kappa_table(const rihvr_params* p) {
    const int Nx = p->width, Ny = p->height;
    const double lambda = (double)p->wavelength;
    const double reso = (double)p->resolution;
    std::vector<double> k((size_t)Nx * Ny);
    for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x) {
            double fx = (double)(((x + Nx / 2) % Nx) - (Nx / 2)) / (double)Nx;
            double fy = (double)(((y + Ny / 2) % Ny) - (Ny / 2)) / (double)Ny;
            double f2 = fx * fx + fy * fy;
            double s = 1 - f2 * (lambda / reso) * (lambda / reso);
            s = (s < 0) ? 0 : s;
            k[(size_t)y * Nx + x] = -2 * kPi * std::sqrt(s) / lambda;
        }
    return k;
}

// The reference passes double phase arguments to sincosf, i.e. rounds them to
// float first (Q2).  The FP64 oracle keeps them in double.
template <class T> inline double round_arg(double v);
template <> inline double round_arg<float>(double v) { return (double)(float)v; }
template <> inline double round_arg<double>(double v) { return v; }

template <class T> inline void sincos_arg(double v, T& s, T& c) {
    double a = round_arg<T>(v);
    s = (T)std::sin(a);
    c = (T)std::cos(a);
}

// ---------------------------------------------------------------------------
// Reductions.  FP32: the reference's 512-thread tree over 1024-element chunks
// (e.g. optical_field.cu:410-446) with the partials summed on the host.
// FP64: plain double sums.
// ---------------------------------------------------------------------------
inline float tree512(float* s) {
    for (int st = 256; st > 0; st >>= 1)
        for (int t = 0; t < st; ++t) s[t] += s[t + st];
    return s[0];
}

// my_summag_kernel (optical_field.cu:410-446) + calcSumMagnitude host float sum (:473-478)
float summag_ref(const C2<float>* d, long n) {
    long nb = (n + 1023) / 1024;
    float sum = 0;
    float s[512];
    auto at = [&](long i) { return i < n ? d[i] : C2<float>{0, 0}; };
    for (long b = 0; b < nb; ++b) {
        for (int t = 0; t < 512; ++t) {
            long i = b * 1024 + t, i2 = i + 512;
            C2<float> a = at(i), c = at(i2);
            float v = 0.0f;
            v += a.x * a.x + c.x * c.x;
            v += a.y * a.y + c.y * c.y;
            s[t] = v;
        }
        sum += tree512(s);
    }
    return sum;
}
double summag(const C2<float>* d, long n) { return (double)summag_ref(d, n); }
double summag(const C2<double>* d, long n) {
    double s = 0;
    for (long i = 0; i < n; ++i) s += d[i].x * d[i].x + d[i].y * d[i].y;
    return s;
}

// calcL1NormComplex (sparse_volume.cu:194-235): double accumulation of float
// moduli of the non-zeros.
double l1_norm(const C2<float>* v, long n) {
    double s = 0;
    for (long i = 0; i < n; ++i)
        if (v[i].x != 0.0f || v[i].y != 0.0f) s += std::sqrt(v[i].x * v[i].x + v[i].y * v[i].y);
    return s;
}
double l1_norm(const C2<double>* v, long n) {
    double s = 0;
    for (long i = 0; i < n; ++i) s += std::sqrt(v[i].x * v[i].x + v[i].y * v[i].y);
    return s;
}

// tv_norm_2d_kernel (sparse_volume.cu:365-390) per pixel: periodic backward
// differences, joint modulus over (re, im) x (x, y).
template <class T> inline T tv_pixel(const C2<T>* d, int x, int y, int W, int H) {
    int xm = (x > 0) ? x - 1 : W - 1;
    int ym = (y > 0) ? y - 1 : H - 1;
    C2<T> c = d[(size_t)y * W + x], a = d[(size_t)y * W + xm], b = d[(size_t)ym * W + x];
    T gxx = a.x - c.x, gyx = b.x - c.x, gxy = a.y - c.y, gyy = b.y - c.y;
    T nrm = gxx * gxx + gxy * gxy + gyx * gyx + gyy * gyy;
    return std::sqrt(nrm);
}

// calcTVNorm2d (sparse_volume.cu:392-446): per plane tv_summation_kernel<512>
// tree (:267-301), partials accumulated on the host in double.
double tv2d_norm(const C2<float>* v, int Nz, int H, int W) {
    const long N = (long)H * W;
    double total = 0;
    std::vector<double> per_plane(Nz);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Nz; ++z) {
        const C2<float>* d = v + (size_t)z * N;
        std::vector<float> pix(N);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) pix[(size_t)y * W + x] = tv_pixel(d, x, y, W, H);
        long nb = (N + 1023) / 1024;
        float s[512];
        double acc = 0;
        for (long b = 0; b < nb; ++b) {
            for (int t = 0; t < 512; ++t) {
                long i = b * 1024 + t, i2 = i + 512;
                float a = 0.0f;
                a += i < N ? pix[i] : 0.0f;
                a += i2 < N ? pix[i2] : 0.0f;
                s[t] = a;
            }
            acc += tree512(s);
        }
        per_plane[z] = acc;
    }
    // the host adds partials plane by plane in z order into one double
    for (int z = 0; z < Nz; ++z) total += per_plane[z];
    return total;
}
double tv2d_norm(const C2<double>* v, int Nz, int H, int W) {
    const long N = (long)H * W;
    double total = 0;
    std::vector<double> per_plane(Nz);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Nz; ++z) {
        const C2<double>* d = v + (size_t)z * N;
        double acc = 0;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) acc += tv_pixel(d, x, y, W, H);
        per_plane[z] = acc;
    }
    for (int z = 0; z < Nz; ++z) total += per_plane[z];
    return total;
}

// tv_norm_kernel (sparse_volume.cu:236-265) per voxel: periodic backward
// differences in x and y, and against the previous plane (plane Nz-1 for plane 0).
template <class T> inline T tv3_pixel(const C2<T>* d, const C2<T>* dp, int x, int y, int W, int H) {
    int xm = (x > 0) ? x - 1 : W - 1;
    int ym = (y > 0) ? y - 1 : H - 1;
    size_t i = (size_t)y * W + x;
    C2<T> c = d[i], a = d[(size_t)y * W + xm], b = d[(size_t)ym * W + x], q = dp[i];
    T gxx = a.x - c.x, gyx = b.x - c.x, gzx = q.x - c.x;
    T gxy = a.y - c.y, gyy = b.y - c.y, gzy = q.y - c.y;
    T nrm = gxx * gxx + gxy * gxy + gyx * gyx + gyy * gyy + gzx * gzx + gzy * gzy;
    return std::sqrt(nrm);
}

// calcTVNorm (sparse_volume.cu:303-363): per plane tv_summation_kernel<512>
// tree, block partials added on the host into one double in z order.
double tv3d_norm(const C2<float>* v, int Nz, int H, int W) {
    const long N = (long)H * W;
    long nb = (N + 1023) / 1024;
    std::vector<std::vector<float>> part(Nz, std::vector<float>(nb));
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Nz; ++z) {
        const C2<float>* d = v + (size_t)z * N;
        const C2<float>* dp = v + (size_t)(z > 0 ? z - 1 : Nz - 1) * N;
        std::vector<float> pix(N);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) pix[(size_t)y * W + x] = tv3_pixel(d, dp, x, y, W, H);
        float s[512];
        for (long b = 0; b < nb; ++b) {
            for (int t = 0; t < 512; ++t) {
                long i = b * 1024 + t, i2 = i + 512;
                float a = 0.0f;
                a += i < N ? pix[i] : 0.0f;
                a += i2 < N ? pix[i2] : 0.0f;
                s[t] = a;
            }
            part[z][b] = tree512(s);
        }
    }
    double total = 0;
    for (int z = 0; z < Nz; ++z)
        for (long b = 0; b < nb; ++b) total += part[z][b];
    return total;
}
double tv3d_norm(const C2<double>* v, int Nz, int H, int W) {
    const long N = (long)H * W;
    double total = 0;
    for (int z = 0; z < Nz; ++z) {
        const C2<double>* d = v + (size_t)z * N;
        const C2<double>* dp = v + (size_t)(z > 0 ? z - 1 : Nz - 1) * N;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) total += tv3_pixel(d, dp, x, y, W, H);
    }
    return total;
}

// Per-plane block partials of sparse_linesearch_1/2_kernel (sparse_compressive_holo.cu:4226-4305)
// and sparse_fistaRestart_kernel (:4393-4432), FP32 semantics.  The host sums
// the partials of all planes in z order into one float (:4348-4352, :4464-4469).
struct LsRestart { double ls1, ls2, restart; };

template <class T> struct PlanePartials { std::vector<T> ls1, ls2, rs; };

void plane_partials(const C2<float>* x1, const C2<float>* x0, const C2<float>* g,
                    const C2<float>* y0, long N, PlanePartials<float>& out, bool want_ls,
                    bool want_restart) {
    long nb = (N + 1023) / 1024;
    float s[512];
    auto at = [&](const C2<float>* p, long i) { return i < N ? p[i] : C2<float>{0, 0}; };
    if (want_ls) {
        out.ls1.resize(nb);
        out.ls2.resize(nb);
        for (long b = 0; b < nb; ++b) {
            for (int t = 0; t < 512; ++t) {
                long i = b * 1024 + t, i2 = i + 512;
                float a = 0.0f;
                C2<float> u1 = at(x1, i), u0 = at(x0, i), gg = at(g, i);
                float tx = (u1.x - u0.x) * gg.x, ty = (u1.y - u0.y) * gg.y;
                a += tx + ty;
                u1 = at(x1, i2); u0 = at(x0, i2); gg = at(g, i2);
                tx = (u1.x - u0.x) * gg.x; ty = (u1.y - u0.y) * gg.y;
                a += tx + ty;
                s[t] = a;
            }
            out.ls1[b] = tree512(s);
            for (int t = 0; t < 512; ++t) {
                long i = b * 1024 + t, i2 = i + 512;
                float a = 0.0f;
                C2<float> u1 = at(x1, i), u0 = at(x0, i);
                float tx = u1.x - u0.x, ty = u1.y - u0.y;
                a += tx * tx + ty * ty;
                u1 = at(x1, i2); u0 = at(x0, i2);
                tx = u1.x - u0.x; ty = u1.y - u0.y;
                a += tx * tx + ty * ty;
                s[t] = a;
            }
            out.ls2[b] = tree512(s);
        }
    }
    if (want_restart) {
        // temp.x is overwritten in the second half (:4411-4412): offsets
        // 0..511 of a chunk add d_re + d_im, offsets 512..1023 add d_im(i2) + d_im(i).
        out.rs.resize(nb);
        for (long b = 0; b < nb; ++b) {
            for (int t = 0; t < 512; ++t) {
                long i = b * 1024 + t, i2 = i + 512;
                C2<float> a0 = at(x0, i), a1 = at(x1, i), ac = at(y0, i);
                float tx = (a0.x - a1.x) * (a1.x - ac.x);
                float ty = (a0.y - a1.y) * (a1.y - ac.y);
                float a = 0.0f;
                a += tx + ty;
                C2<float> b0 = at(x0, i2), b1 = at(x1, i2), bc = at(y0, i2);
                tx = (b0.x - b1.x) * (b1.x - bc.x);
                tx = (b0.y - b1.y) * (b1.y - bc.y);
                a += tx + ty;
                s[t] = a;
            }
            out.rs[b] = tree512(s);
        }
    }
}

void plane_partials(const C2<double>* x1, const C2<double>* x0, const C2<double>* g,
                    const C2<double>* y0, long N, PlanePartials<double>& out, bool want_ls,
                    bool want_restart) {
    double l1 = 0, l2 = 0, rs = 0;
    for (long i = 0; i < N; ++i) {
        if (want_ls) {
            double dx = x1[i].x - x0[i].x, dy = x1[i].y - x0[i].y;
            l1 += dx * g[i].x + dy * g[i].y;
            l2 += dx * dx + dy * dy;
        }
        if (want_restart) {
            // same quirky weights as the reference kernel, in double
            long off = i % 1024;
            double dre = (x0[i].x - x1[i].x) * (x1[i].x - y0[i].x);
            double dim = (x0[i].y - x1[i].y) * (x1[i].y - y0[i].y);
            rs += off < 512 ? dre + 2.0 * dim : dim;
        }
    }
    out.ls1.assign(1, l1);
    out.ls2.assign(1, l2);
    out.rs.assign(1, rs);
}

template <class T> struct HostSum;
template <> struct HostSum<float> { using type = float; };
template <> struct HostSum<double> { using type = double; };

// ---------------------------------------------------------------------------
// The solver state for one context
// ---------------------------------------------------------------------------
template <class T> struct Solver {
    rihvr_params p;
    int W, H, Nz;
    long N;
    std::vector<float> z;
    std::vector<double> kap;
    FFT2<T> fft;
    int nthreads;

    explicit Solver(const rihvr_params* prm)
        : p(*prm), W(prm->width), H(prm->height), Nz(prm->num_planes),
          N((long)prm->width * prm->height), z(plane_list(prm)), kap(kappa_table(prm)),
          fft(prm->height, prm->width) {
#ifdef _OPENMP
        nthreads = omp_get_max_threads();
#else
        nthreads = 1;
#endif
    }

    // A: SparseCompressiveHolo::backPropagate (sparse_compressive_holo.cu:1935-2077)
    // with sparse_phase_scale_kernel (:1842-1861), sparse_rs_phase_mult_kernel
    // (:1863-1888), sparse_add_plane_kernel (:1890-1903) and
    // back_propagate_finish_kernel (:1918-1933).
    void forward(const C2<T>* vol, C2<T>* est) const {
        std::vector<C2<T>> S(N, C2<T>{0, 0});
        const double lambda = (double)p.wavelength;
        const T invN = (T)N;
        int chunk = std::max(1, nthreads);
        std::vector<C2<T>> bufs((size_t)chunk * N), work((size_t)chunk * N);
        std::vector<char> used(chunk);
        for (int z0 = 0; z0 < Nz; z0 += chunk) {
            int zc = std::min(chunk, Nz - z0);
#pragma omp parallel for schedule(static, 1)
            for (int c = 0; c < zc; ++c) {
                int zi = z0 + c;
                const C2<T>* x = vol + (size_t)zi * N;
                bool nz = false;
                for (long i = 0; i < N && !nz; ++i) nz = (x[i].x != (T)0 || x[i].y != (T)0);
                used[c] = nz;
                if (!nz) continue;  // planes with nnz == 0 are skipped (:1970)
                C2<T>* b = bufs.data() + (size_t)c * N;
                double zz = -(double)z[zi];
                T py, px;
                sincos_arg<T>(2 * kPi * zz / lambda, py, px);
                for (long i = 0; i < N; ++i) {
                    C2<T> d = x[i];
                    b[i] = {d.x * px - d.y * py, d.y * px + d.x * py};
                }
                fft.exec(b, work.data() + (size_t)c * N, -1);
                for (long i = 0; i < N; ++i) {
                    T hy, hx;
                    sincos_arg<T>(zz * kap[i], hy, hx);
                    C2<T> in = b[i];
                    T re = hx * in.x - hy * in.y;
                    T im = hy * in.x + hx * in.y;
                    b[i] = {(T)((double)re / (double)N), (T)((double)im / (double)N)};
                }
            }
            (void)invN;
            for (int c = 0; c < zc; ++c) {
                if (!used[c]) continue;
                const C2<T>* b = bufs.data() + (size_t)c * N;
                for (long i = 0; i < N; ++i) { S[i].x += b[i].x; S[i].y += b[i].y; }
            }
        }
        fft.exec(S.data(), work.data(), +1);
        for (long i = 0; i < N; ++i) est[i] = {S[i].x, (T)0};
    }

    // R_hat = FFT(r): SparseGradient::reconstruct (sparse_gradient.cu:47-75)
    void spectrum(const C2<T>* r, C2<T>* Rh) const {
        std::memcpy(Rh, r, sizeof(C2<T>) * N);
        std::vector<C2<T>> work(N);
        fft.exec(Rh, work.data(), -1);
    }

    // A^T for every plane: SparseGradient::getPlane (sparse_gradient.cu:125-165)
    // with sparse_rs_phase_mult_kernel2 (:77-101) and sparse_phase_scale_kernel2 (:103-123).
    void adjoint_from_spectrum(const C2<T>* Rh, C2<T>* G) const {
        const double lambda = (double)p.wavelength;
        std::vector<C2<T>> work((size_t)std::max(1, nthreads) * N);
#pragma omp parallel for schedule(static, 1)
        for (int zi = 0; zi < Nz; ++zi) {
#ifdef _OPENMP
            int tid = omp_get_thread_num();
#else
            int tid = 0;
#endif
            C2<T>* g = G + (size_t)zi * N;
            double zz = (double)z[zi];
            for (long i = 0; i < N; ++i) {
                T hy, hx;
                sincos_arg<T>(zz * kap[i], hy, hx);
                C2<T> in = Rh[i];
                g[i] = {(hx * in.x - hy * in.y) / (T)N, (hy * in.x + hx * in.y) / (T)N};
            }
            fft.exec(g, work.data() + (size_t)tid * N, +1);
            T py, px;
            sincos_arg<T>(2 * kPi * zz / lambda, py, px);
            for (long i = 0; i < N; ++i) {
                C2<T> d = g[i];
                g[i] = {d.x * px - d.y * py, d.y * px + d.x * py};
            }
        }
    }

    // block_soft_threshold_kernel (sparse_compressive_holo.cu:3065-3110)
    static inline T soft1(T v, T thr) {
        T a = std::fabs(v);
        T y = std::max(a - thr, (T)0);
        y = y / (y + thr);
        return y * v;
    }

    // proximalFusedLasso2d for one plane (:3652-3732) with sparse_get_xhat_kernel
    // (:3478-3489), tv_denoised_estimate_2d_kernel (:3491-3513),
    // tv_gradient_projection_2d_kernel (:3539-3597), block_soft_threshold_kernel.
    void prox_fl2d_plane(const C2<T>* x0, const C2<T>* g, double t, C2<T>* out,
                         std::vector<C2<T>>& xh, std::vector<C2<T>>& gx,
                         std::vector<C2<T>>& gy, std::vector<C2<T>>& yh) const {
        const double regTV = (double)p.regularization_TV_param;
        const double reg = (double)p.regularization_param;
        const double scale = 1 / (4 * 3 * t * regTV);
        const double mu = t * regTV;
        for (long i = 0; i < N; ++i)
            xh[i] = {(T)((double)x0[i].x - t * (double)g[i].x),
                     (T)((double)x0[i].y - t * (double)g[i].y)};
        std::fill(gx.begin(), gx.end(), C2<T>{0, 0});
        std::fill(gy.begin(), gy.end(), C2<T>{0, 0});
        auto estimate = [&]() {
            for (int y = 0; y < H; ++y) {
                int yp = (y < H - 1) ? y + 1 : 0;
                for (int x = 0; x < W; ++x) {
                    int xp = (x < W - 1) ? x + 1 : 0;
                    size_t idx = (size_t)y * W + x;
                    size_t ixp = (size_t)y * W + xp, iyp = (size_t)yp * W + x;
                    T dx = gx[ixp].x - gx[idx].x + gy[iyp].x - gy[idx].x;
                    T dy = gx[ixp].y - gx[idx].y + gy[iyp].y - gy[idx].y;
                    yh[idx] = {(T)((double)xh[idx].x - mu * (double)dx),
                               (T)((double)xh[idx].y - mu * (double)dy)};
                }
            }
        };
        for (int it = 0; it < p.num_tv_iterations; ++it) {
            estimate();
            for (int y = 0; y < H; ++y) {
                int ym = (y > 0) ? y - 1 : H - 1;
                for (int x = 0; x < W; ++x) {
                    int xm = (x > 0) ? x - 1 : W - 1;
                    size_t idx = (size_t)y * W + x;
                    size_t ixm = (size_t)y * W + xm, iym = (size_t)ym * W + x;
                    T gxx = yh[ixm].x - yh[idx].x;
                    T gyx = yh[iym].x - yh[idx].x;
                    T gxy = yh[ixm].y - yh[idx].y;
                    T gyy = yh[iym].y - yh[idx].y;
                    gxx = (T)((double)gx[idx].x + scale * (double)gxx);
                    gyx = (T)((double)gy[idx].x + scale * (double)gyx);
                    gxy = (T)((double)gx[idx].y + scale * (double)gxy);
                    gyy = (T)((double)gy[idx].y + scale * (double)gyy);
                    T nrm = gxx * gxx + gxy * gxy + gyx * gyx + gyy * gyy;
                    nrm = (T)std::max((double)std::sqrt(nrm), 1.0);
                    gx[idx] = {gxx / nrm, gxy / nrm};
                    gy[idx] = {gyx / nrm, gyy / nrm};
                }
            }
        }
        estimate();
        if (reg > 0.0) {
            const T thr = (T)(t * reg);
            for (long i = 0; i < N; ++i) out[i] = {soft1(yh[i].x, thr), soft1(yh[i].y, thr)};
        } else {
            for (long i = 0; i < N; ++i) out[i] = yh[i];
        }
    }

    // proximalL1 for one plane (:3274-3392) = block_soft_threshold_FASTA_kernel (:3112-3157)
    void prox_l1_plane(const C2<T>* x0, const C2<T>* g, double t, C2<T>* out) const {
        const double reg = (double)p.regularization_param;
        const T thr = (T)(t * reg);
        for (long i = 0; i < N; ++i) {
            T inx = (T)((double)x0[i].x - t * (double)g[i].x);
            T iny = (T)((double)x0[i].y - t * (double)g[i].y);
            out[i] = {soft1(inx, thr), soft1(iny, thr)};
        }
    }

    // proximalFusedLasso, mode 6 (:3422-3475): block_soft_threshold_FASTA_kernel at
    // regularization/2, totalVariationProjection (:3867-3888) over the whole volume with
    // tv_denoised_estimate_kernel (:3890-3915) and tv_gradient_projection_kernel
    // (:3964-3997) -- z-periodic, normalizer forced to 1 (:3991) -- then
    // block_soft_threshold_kernel at regularization/2.
    void prox_fl3d(const C2<T>* x0, const C2<T>* G, double t, C2<T>* out) const {
        const long V = (long)Nz * N;
        const double reg = (double)p.regularization_param;
        const double regTV = (double)p.regularization_TV_param;
        const double scale = 1 / (4 * 3 * t * regTV);
        const double mu = t * regTV;
        const T thr = (T)(t * (reg / 2.0));
        std::vector<C2<T>> x1(V), y(V), gx(V, C2<T>{0, 0}), gy(V, C2<T>{0, 0}), gz(V, C2<T>{0, 0});
#pragma omp parallel for schedule(static)
        for (long i = 0; i < V; ++i) {
            T inx = (T)((double)x0[i].x - t * (double)G[i].x);
            T iny = (T)((double)x0[i].y - t * (double)G[i].y);
            x1[i] = {soft1(inx, thr), soft1(iny, thr)};
        }
        auto estimate = [&]() {
#pragma omp parallel for schedule(static)
            for (int zi = 0; zi < Nz; ++zi) {
                const size_t o = (size_t)zi * N, on = (size_t)(zi < Nz - 1 ? zi + 1 : 0) * N;
                for (int yy = 0; yy < H; ++yy) {
                    int yp = (yy < H - 1) ? yy + 1 : 0;
                    for (int xx = 0; xx < W; ++xx) {
                        int xp = (xx < W - 1) ? xx + 1 : 0;
                        size_t idx = (size_t)yy * W + xx;
                        size_t ixp = (size_t)yy * W + xp, iyp = (size_t)yp * W + xx;
                        T dx = gx[o + ixp].x - gx[o + idx].x + gy[o + iyp].x - gy[o + idx].x + gz[on + idx].x -
                               gz[o + idx].x;
                        T dy = gx[o + ixp].y - gx[o + idx].y + gy[o + iyp].y - gy[o + idx].y + gz[on + idx].y -
                               gz[o + idx].y;
                        y[o + idx] = {(T)((double)x1[o + idx].x - mu * (double)dx),
                                      (T)((double)x1[o + idx].y - mu * (double)dy)};
                    }
                }
            }
        };
        for (int it = 0; it < p.num_tv_iterations; ++it) {
            estimate();
#pragma omp parallel for schedule(static)
            for (int zi = 0; zi < Nz; ++zi) {
                const size_t o = (size_t)zi * N, op = (size_t)(zi > 0 ? zi - 1 : Nz - 1) * N;
                for (int yy = 0; yy < H; ++yy) {
                    int ym = (yy > 0) ? yy - 1 : H - 1;
                    for (int xx = 0; xx < W; ++xx) {
                        int xm = (xx > 0) ? xx - 1 : W - 1;
                        size_t idx = (size_t)yy * W + xx;
                        size_t ixm = (size_t)yy * W + xm, iym = (size_t)ym * W + xx;
                        const C2<T> c = y[o + idx];
                        T a[6] = {y[o + ixm].x - c.x, y[o + iym].x - c.x, y[op + idx].x - c.x,
                                  y[o + ixm].y - c.y, y[o + iym].y - c.y, y[op + idx].y - c.y};
                        gx[o + idx] = {(T)((double)gx[o + idx].x + scale * (double)a[0]),
                                       (T)((double)gx[o + idx].y + scale * (double)a[3])};
                        gy[o + idx] = {(T)((double)gy[o + idx].x + scale * (double)a[1]),
                                       (T)((double)gy[o + idx].y + scale * (double)a[4])};
                        gz[o + idx] = {(T)((double)gz[o + idx].x + scale * (double)a[2]),
                                       (T)((double)gz[o + idx].y + scale * (double)a[5])};
                    }
                }
            }
        }
        estimate();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < V; ++i) out[i] = {soft1(y[i].x, thr), soft1(y[i].y, thr)};
    }

    // denoise (:2948-2995) dispatch, every plane
    int prox(int mode, const C2<T>* x0, const C2<T>* G, double t, C2<T>* out) const {
        if (mode == RIHVR_MODE_FASTA_L1) {
#pragma omp parallel for schedule(static, 1)
            for (int zi = 0; zi < Nz; ++zi)
                prox_l1_plane(x0 + (size_t)zi * N, G + (size_t)zi * N, t, out + (size_t)zi * N);
            return RIHVR_OK;
        }
        if (mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D) {
#pragma omp parallel
            {
                std::vector<C2<T>> xh(N), gx(N), gy(N), yh(N);
#pragma omp for schedule(static, 1)
                for (int zi = 0; zi < Nz; ++zi)
                    prox_fl2d_plane(x0 + (size_t)zi * N, G + (size_t)zi * N, t,
                                    out + (size_t)zi * N, xh, gx, gy, yh);
            }
            return RIHVR_OK;
        }
        if (mode == RIHVR_MODE_FASTA_FUSED_LASSO) {
            prox_fl3d(x0, G, t, out);
            return RIHVR_OK;
        }
        return RIHVR_E_INVALID_ARGUMENT;
    }

    // calcObjectiveFunction (:2838-2899)
    double objective(int mode, const C2<T>* r, const C2<T>* x) const {
        double resid_norm = summag(r, N);
        double est_norm = l1_norm(x, (long)Nz * N);
        double obj = 0.5 * resid_norm + (double)p.regularization_param * est_norm;
        if (mode == RIHVR_MODE_FASTA_FUSED_LASSO_2D)
            obj = obj + (double)p.regularization_TV_param * tv2d_norm(x, Nz, H, W);
        else if (mode == RIHVR_MODE_FASTA_FUSED_LASSO)
            obj = obj + (double)p.regularization_TV_param * tv3d_norm(x, Nz, H, W);
        return obj;
    }

    // calcLineSearchLimit (:4307-4390) and fistaRestart (:4434-4482): host
    // accumulators sum1/sum2 are float for the FP32 semantics.
    void ls_restart(const C2<T>* x1, const C2<T>* x0, const C2<T>* G, const C2<T>* y0,
                    double t, bool want_ls, bool want_restart, double* lim,
                    double* rsum) const {
        std::vector<PlanePartials<T>> parts(Nz);
#pragma omp parallel for schedule(static, 1)
        for (int zi = 0; zi < Nz; ++zi) {
            size_t o = (size_t)zi * N;
            plane_partials(x1 + o, x0 + o, G ? G + o : nullptr, y0 ? y0 + o : nullptr, N,
                           parts[zi], want_ls, want_restart);
        }
        using H_t = typename HostSum<T>::type;
        H_t s1 = 0, s2 = 0, sr = 0;
        for (int zi = 0; zi < Nz; ++zi) {
            if (want_ls) {
                for (T v : parts[zi].ls1) s1 += v;
                for (T v : parts[zi].ls2) s2 += v;
            }
            if (want_restart)
                for (T v : parts[zi].rs) sr += v;
        }
        if (want_ls) *lim = (double)s1 + (double)s2 / (2 * t);
        if (want_restart) *rsum = (double)sr;
    }

    // sparse_fistaUpdate_kernel (:4484-4493): x1 = y1 + step*(y1 - y0)
    static void momentum(C2<T>* out, const C2<T>* y0, const C2<T>* y1, double s, long n) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            T dx = y1[i].x - y0[i].x, dy = y1[i].y - y0[i].y;
            out[i] = {(T)((double)y1[i].x + s * (double)dx), (T)((double)y1[i].y + s * (double)dy)};
        }
    }

    // inverseReconstruct (:173-633), modes 5, 6 and 7
    int solve(const C2<T>* meas, const C2<T>* warm, C2<T>* out, C2<T>* est_out,
              C2<T>* res_out, rihvr_stats* st) {
        int mode = p.mode;
        if (mode != RIHVR_MODE_FASTA_L1 && mode != RIHVR_MODE_FASTA_FUSED_LASSO &&
            mode != RIHVR_MODE_FASTA_FUSED_LASSO_2D)
            return RIHVR_E_INVALID_ARGUMENT;  // :178-189
        if (p.regularization_TV_param == 0.0f) mode = RIHVR_MODE_FASTA_L1;  // :192-195, any mode
        for (long i = 0; i < N; ++i)
            if (!std::isfinite((double)meas[i].x) || !std::isfinite((double)meas[i].y))
                return RIHVR_E_INVALID_DATA;
        const long V = (long)Nz * N;
        double t = 1.0 / (double)Nz;  // :247
        std::vector<C2<T>> x1(V), x0(V), y0(V), G(V);
        if (warm) std::memcpy(x1.data(), warm, sizeof(C2<T>) * V);
        else std::fill(x1.begin(), x1.end(), C2<T>{0, 0});
        x0 = x1;
        y0 = x1;
        double alpha1 = 1.0;
        std::vector<C2<T>> est(N), r(N), Rh(N), cur(N, C2<T>{0, 0}), prev(N, C2<T>{0, 0});
        auto residual = [&]() {
            for (long i = 0; i < N; ++i) r[i] = {est[i].x - meas[i].x, est[i].y - meas[i].y};
        };
        forward(x1.data(), est.data());
        residual();
        double f = objective(mode, r.data(), x1.data());
        double f_resid = 0.5 * summag(r.data(), N);
        double prev_f_resid = f_resid;
        spectrum(r.data(), Rh.data());
        adjoint_from_spectrum(Rh.data(), G.data());
        int total_bt = 0;
        if (st) { st->f_initial = f; st->iterations = 0; }
        // RIHVR_ORACLE_PROGRESS=k: a line on stderr every k iterations (long solves)
        const char* pe = std::getenv("RIHVR_ORACLE_PROGRESS");
        const int progress = pe ? std::atoi(pe) : 0;
        for (int it = 0; it < p.num_inverse_iterations; ++it) {
            if (progress > 0 && it > 0 && it % progress == 0) {
                std::fprintf(stderr, "oracle solve: %d of %d iterations done\n", it, p.num_inverse_iterations);
                std::fflush(stderr);
            }
            x0 = x1;
            double alpha0 = alpha1;
            prox(mode, x0.data(), G.data(), t, x1.data());
            forward(x1.data(), est.data());
            residual();
            double lim = 0, rs = 0;
            ls_restart(x1.data(), x0.data(), G.data(), nullptr, t, true, false, &lim, nullptr);
            f_resid = 0.5 * summag(r.data(), N);
            double ratio = f_resid / (prev_f_resid + lim);
            int bt = 0;
            while (((ratio > 1.0) || (ratio < 0)) && (bt < 2)) {
                t *= 0.5;
                prox(mode, x0.data(), G.data(), t, x1.data());
                forward(x1.data(), est.data());
                residual();
                ls_restart(x1.data(), x0.data(), G.data(), nullptr, t, true, false, &lim, nullptr);
                f_resid = 0.5 * summag(r.data(), N);
                ratio = f_resid / (prev_f_resid + lim);
                bt++;
                total_bt++;
            }
            ls_restart(x1.data(), x0.data(), nullptr, y0.data(), t, false, true, nullptr, &rs);
            bool restart = rs > 0;
            if (restart) alpha0 = 1;
            alpha1 = (1.0 + std::sqrt(1.0 + 4.0 * alpha0 * alpha0)) / 2.0;
            double step = (alpha0 - 1) / alpha1;
            // y1 = x1; x1 = y1 + step*(y1 - y0); y0 = y1   (:478-489)
            std::vector<C2<T>>& y1 = x0;  // x0 is dead until the next iteration
            y1 = x1;
            momentum(x1.data(), y0.data(), y1.data(), step, V);
            std::swap(y0, y1);
            // hologram momentum by linearity (:494-498, :4599-4642)
            prev = cur;
            cur = est;
            momentum(est.data(), prev.data(), cur.data(), step, N);
            prev = cur;
            residual();  // calcResidualFrom (:4644-4742)
            spectrum(r.data(), Rh.data());
            adjoint_from_spectrum(Rh.data(), G.data());
            f = objective(mode, r.data(), x1.data());
            prev_f_resid = 0.5 * summag(r.data(), N);
            if (st) {
                st->iterations = it + 1;
                if (st->f) st->f[it] = f;
                if (st->f_resid) st->f_resid[it] = prev_f_resid;
                if (st->ratio) st->ratio[it] = ratio;
                if (st->lim) st->lim[it] = lim;
                if (st->backtracks) st->backtracks[it] = bt;
                if (st->restarts) st->restarts[it] = restart ? 1 : 0;
                if (st->restart_sum) st->restart_sum[it] = rs;
            }
        }
        std::memcpy(out, x1.data(), sizeof(C2<T>) * V);
        if (est_out) std::memcpy(est_out, est.data(), sizeof(C2<T>) * N);
        if (res_out) std::memcpy(res_out, r.data(), sizeof(C2<T>) * N);
        if (st) {
            long nnz = 0;
            for (long i = 0; i < V; ++i) nnz += (x1[i].x != (T)0 || x1[i].y != (T)0);
            st->total_backtracks = total_bt;
            st->stepsize = t;
            st->f_final = f;
            st->nnz = nnz;
            st->sparsity = (double)nnz / (double)V;
        }
        return RIHVR_OK;
    }
};

int check_params(const rihvr_params* p) {
    if (!p) return RIHVR_E_INVALID_ARGUMENT;
    if (p->width <= 0 || p->height <= 0 || p->num_planes <= 0) return RIHVR_E_INVALID_ARGUMENT;
    if (!(p->wavelength > 0) || !(p->resolution > 0)) return RIHVR_E_INVALID_ARGUMENT;
    if (p->num_inverse_iterations < 0 || p->num_tv_iterations < 0) return RIHVR_E_INVALID_ARGUMENT;
    return RIHVR_OK;
}

template <class T> int forward_impl(const rihvr_params* p, const T* vol, T* holo) {
    int rc = check_params(p);
    if (rc) return rc;
    Solver<T> s(p);
    s.forward((const C2<T>*)vol, (C2<T>*)holo);
    return RIHVR_OK;
}
template <class T> int adjoint_impl(const rihvr_params* p, const T* holo, T* vol) {
    int rc = check_params(p);
    if (rc) return rc;
    Solver<T> s(p);
    std::vector<C2<T>> Rh(s.N);
    s.spectrum((const C2<T>*)holo, Rh.data());
    s.adjoint_from_spectrum(Rh.data(), (C2<T>*)vol);
    return RIHVR_OK;
}
template <class T>
int prox_impl(const rihvr_params* p, const T* x0, const T* g, double t, T* out) {
    int rc = check_params(p);
    if (rc) return rc;
    Solver<T> s(p);
    return s.prox(p->mode, (const C2<T>*)x0, (const C2<T>*)g, t, (C2<T>*)out);
}
template <class T>
int solve_impl(const rihvr_params* p, const T* holo, const T* warm, T* out, T* est, T* res,
               rihvr_stats* st) {
    int rc = check_params(p);
    if (rc) return rc;
    Solver<T> s(p);
    return s.solve((const C2<T>*)holo, (const C2<T>*)warm, (C2<T>*)out, (C2<T>*)est,
                   (C2<T>*)res, st);
}
template <class T> int fft2_impl(T* d, int H, int W, int sign) {
    if (H <= 0 || W <= 0) return RIHVR_E_INVALID_ARGUMENT;
    FFT2<T> f(H, W);
    std::vector<C2<T>> work((size_t)H * W);
    f.exec((C2<T>*)d, work.data(), sign);
    return RIHVR_OK;
}

}  // namespace

extern "C" {

int ro_version(void) { return 1; }

void ro_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int ro_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int ro_fft2_f32(float* d, int H, int W, int sign) { return fft2_impl(d, H, W, sign); }
int ro_fft2_f64(double* d, int H, int W, int sign) { return fft2_impl(d, H, W, sign); }

int ro_kappa(const rihvr_params* p, double* out) {
    int rc = check_params(p);
    if (rc) return rc;
    auto k = kappa_table(p);
    std::memcpy(out, k.data(), sizeof(double) * k.size());
    return RIHVR_OK;
}
int ro_plane_list(const rihvr_params* p, float* out) {
    int rc = check_params(p);
    if (rc) return rc;
    auto z = plane_list(p);
    std::memcpy(out, z.data(), sizeof(float) * z.size());
    return RIHVR_OK;
}

int ro_forward_f32(const rihvr_params* p, const float* v, float* h) { return forward_impl(p, v, h); }
int ro_forward_f64(const rihvr_params* p, const double* v, double* h) { return forward_impl(p, v, h); }
int ro_adjoint_f32(const rihvr_params* p, const float* h, float* v) { return adjoint_impl(p, h, v); }
int ro_adjoint_f64(const rihvr_params* p, const double* h, double* v) { return adjoint_impl(p, h, v); }
int ro_prox_f32(const rihvr_params* p, const float* x0, const float* g, double t, float* o) {
    return prox_impl(p, x0, g, t, o);
}
int ro_prox_f64(const rihvr_params* p, const double* x0, const double* g, double t, double* o) {
    return prox_impl(p, x0, g, t, o);
}

double ro_summag_f32(const float* d, long n) { return summag((const C2<float>*)d, n); }
double ro_l1_f32(const float* v, long n) { return l1_norm((const C2<float>*)v, n); }
double ro_tv2d_f32(const float* v, int Nz, int H, int W) {
    return tv2d_norm((const C2<float>*)v, Nz, H, W);
}
double ro_tv3d_f32(const float* v, int Nz, int H, int W) {
    return tv3d_norm((const C2<float>*)v, Nz, H, W);
}
double ro_restart_sum_f32(const float* x0, const float* x1, const float* y0, long n_plane, int Nz) {
    float s = 0;
    for (int z = 0; z < Nz; ++z) {
        PlanePartials<float> pp;
        size_t o = (size_t)z * n_plane;
        plane_partials((const C2<float>*)x1 + o, (const C2<float>*)x0 + o, nullptr,
                       (const C2<float>*)y0 + o, n_plane, pp, false, true);
        for (float v : pp.rs) s += v;
    }
    return (double)s;
}

int ro_solve_f32(const rihvr_params* p, const float* holo, const float* warm, float* out,
                 float* est, float* res, rihvr_stats* st) {
    return solve_impl(p, holo, warm, out, est, res, st);
}
int ro_solve_f64(const rihvr_params* p, const double* holo, const double* warm, double* out,
                 double* est, double* res, rihvr_stats* st) {
    return solve_impl(p, holo, warm, out, est, res, st);
}

}  // extern "C"
