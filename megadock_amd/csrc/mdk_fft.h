// Device-side building blocks of the batched 3D FFT: a mixed-radix Stockham
// transform of many lines held in LDS, laid out element-major ([elem][line], line
// stride LS) so that lanes of a wavefront touch consecutive LDS words.
//
// Stage with radix R and span Ns (product of the radices already applied):
//   for j < N/R:  k = j mod Ns,
//     v_r = in[j + r N/R] * w^(r k),  w = exp(-+ 2 pi i / (Ns R))
//     out[(j - k) R + k + m Ns] = sum_r v_r exp(-+ 2 pi i r m / R)
// which leaves the result in natural order after the last stage.
#pragma once
#include <hip/hip_runtime.h>

#define MDK_MAX_STAGES 16

struct FftPlan {
    int n;                        // transform length
    int nst;                      // number of stages
    int radix[MDK_MAX_STAGES];    // radix of each stage
    int twoff[MDK_MAX_STAGES];    // offset of each stage's (R-1)*Ns twiddles
    int ntw;                      // total twiddles
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// cos/sin(2 pi k / R) for k = 0..R-1
template <int R> struct DftConst;
template <> struct DftConst<3> {
    static __device__ __forceinline__ float c(int k) {
        const float t[3] = {1.0f, -5.0e-01f, -5.0e-01f};
        return t[k];
    }
    static __device__ __forceinline__ float s(int k) {
        const float t[3] = {0.0f, 8.66025403784438597e-01f, -8.66025403784438597e-01f};
        return t[k];
    }
};
template <> struct DftConst<5> {
    static __device__ __forceinline__ float c(int k) {
        const float t[5] = {1.0f, 3.09016994374947451e-01f, -8.09016994374947451e-01f,
                            -8.09016994374947451e-01f, 3.09016994374947451e-01f};
        return t[k];
    }
    static __device__ __forceinline__ float s(int k) {
        const float t[5] = {0.0f, 9.51056516295153531e-01f, 5.87785252292473137e-01f,
                            -5.87785252292473137e-01f, -9.51056516295153531e-01f};
        return t[k];
    }
};
template <> struct DftConst<7> {
    static __device__ __forceinline__ float c(int k) {
        const float t[7] = {1.0f, 6.23489801858733530e-01f, -2.22520933956314404e-01f,
                            -9.00968867902419126e-01f, -9.00968867902419126e-01f,
                            -2.22520933956314404e-01f, 6.23489801858733530e-01f};
        return t[k];
    }
    static __device__ __forceinline__ float s(int k) {
        const float t[7] = {0.0f, 7.81831482468029809e-01f, 9.74927912181823607e-01f,
                            4.33883739117558120e-01f, -4.33883739117558120e-01f,
                            -9.74927912181823607e-01f, -7.81831482468029809e-01f};
        return t[k];
    }
};
template <> struct DftConst<11> {
    static __device__ __forceinline__ float c(int k) {
        const float t[11] = {1.0f, 8.41253532831181169e-01f, 4.15415013001886425e-01f,
                             -1.42314838273285140e-01f, -6.54860733945285065e-01f,
                             -9.59492973614497390e-01f, -9.59492973614497390e-01f,
                             -6.54860733945285065e-01f, -1.42314838273285140e-01f,
                             4.15415013001886425e-01f, 8.41253532831181169e-01f};
        return t[k];
    }
    static __device__ __forceinline__ float s(int k) {
        const float t[11] = {0.0f, 5.40640817455597582e-01f, 9.09631995354518371e-01f,
                             9.89821441880932732e-01f, 7.55749574354258283e-01f,
                             2.81732556841429697e-01f, -2.81732556841429697e-01f,
                             -7.55749574354258283e-01f, -9.89821441880932732e-01f,
                             -9.09631995354518371e-01f, -5.40640817455597582e-01f};
        return t[k];
    }
};
template <> struct DftConst<13> {
    static __device__ __forceinline__ float c(int k) {
        const float t[13] = {1.0f, 8.85456025653209896e-01f, 5.68064746731155804e-01f,
                             1.20536680255323012e-01f, -3.54604887042535626e-01f,
                             -7.48510748171101099e-01f, -9.70941817426052027e-01f,
                             -9.70941817426052027e-01f, -7.48510748171101099e-01f,
                             -3.54604887042535626e-01f, 1.20536680255323012e-01f,
                             5.68064746731155804e-01f, 8.85456025653209896e-01f};
        return t[k];
    }
    static __device__ __forceinline__ float s(int k) {
        const float t[13] = {0.0f, 4.64723172043768544e-01f, 8.22983865893656400e-01f,
                             9.92708874098053959e-01f, 9.35016242685414803e-01f,
                             6.63122658240795240e-01f, 2.39315664287557714e-01f,
                             -2.39315664287557714e-01f, -6.63122658240795240e-01f,
                             -9.35016242685414803e-01f, -9.92708874098053959e-01f,
                             -8.22983865893656400e-01f, -4.64723172043768544e-01f};
        return t[k];
    }
};

// In-register R-point DFT, forward (exp(-2 pi i ...)) unless INV.
template <int R, bool INV> struct Dft;

template <bool INV> struct Dft<2, INV> {
    static __device__ __forceinline__ void run(float2 *v) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV> struct Dft<4, INV> {
    static __device__ __forceinline__ void run(float2 *v) {
        const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        const float2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
        // forward: b1 * (-i); inverse: b1 * (+i)
        const float2 t = INV ? make_float2(-b1.y, b1.x) : make_float2(b1.y, -b1.x);
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, t);
        v[3] = csub(a1, t);
    }
};

template <bool INV> struct Dft<8, INV> {
    static __device__ __forceinline__ void run(float2 *v) {
        // two radix-4 on even/odd, twiddle by w8^k, radix-2 combine
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV>::run(e);
        Dft<4, INV>::run(o);
        const float h = 7.07106781186547524e-01f;
        // o1 *= w8, o2 *= w8^2, o3 *= w8^3 (w8 = exp(-+ i pi / 4))
        if (!INV) {
            o[1] = make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
            o[2] = make_float2(o[2].y, -o[2].x);
            o[3] = make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
        } else {
            o[1] = make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y));
            o[2] = make_float2(-o[2].y, o[2].x);
            o[3] = make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y));
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};

// odd prime R: pair t with R-t
template <int R, bool INV> struct Dft {
    static __device__ __forceinline__ void run(float2 *v) {
        constexpr int H = (R - 1) / 2;
        float2 s[H], d[H];
        float2 y0 = v[0];
#pragma unroll
        for (int t = 1; t <= H; t++) {
            s[t - 1] = cadd(v[t], v[R - t]);
            d[t - 1] = csub(v[t], v[R - t]);
            y0 = cadd(y0, s[t - 1]);
        }
        float2 out[R];
        out[0] = y0;
#pragma unroll
        for (int m = 1; m <= H; m++) {
            float2 re = v[0], im = make_float2(0.f, 0.f);
#pragma unroll
            for (int t = 1; t <= H; t++) {
                const int e = (t * m) % R;
                const float c = DftConst<R>::c(e), sn = DftConst<R>::s(e);
                re.x += s[t - 1].x * c;
                re.y += s[t - 1].y * c;
                im.x += d[t - 1].x * sn;
                im.y += d[t - 1].y * sn;
            }
            // forward: y[m] = re - i*im, y[R-m] = re + i*im
            const float2 mi = make_float2(im.y, -im.x);  // -i * im
            if (!INV) {
                out[m] = cadd(re, mi);
                out[R - m] = csub(re, mi);
            } else {
                out[m] = csub(re, mi);
                out[R - m] = cadd(re, mi);
            }
        }
#pragma unroll
        for (int m = 0; m < R; m++) v[m] = out[m];
    }
};

// One Stockham stage over the lines of a workgroup.  Thread (line, jt) handles
// butterflies j = jt, jt + tpl, ...  `tw` holds this stage's (R-1)*Ns twiddles
// exp(-2 pi i r k / (Ns R)), index (r-1)*Ns + k.
template <int R, bool INV>
__device__ __forceinline__ void fft_stage(const float2 *__restrict__ in, float2 *__restrict__ out,
                                          const float2 *__restrict__ tw, int N, int Ns, int LS,
                                          int line, int jt, int tpl) {
    const int nr = N / R;
    for (int j = jt; j < nr; j += tpl) {
        const int k = j % Ns;
        float2 v[R];
        v[0] = in[j * LS + line];
#pragma unroll
        for (int r = 1; r < R; r++) {
            float2 w = tw[(r - 1) * Ns + k];
            if (INV) w.y = -w.y;
            v[r] = cmul(in[(j + r * nr) * LS + line], w);
        }
        Dft<R, INV>::run(v);
        const int d = (j - k) * R + k;
#pragma unroll
        for (int m = 0; m < R; m++) out[(d + m * Ns) * LS + line] = v[m];
    }
}

// Full transform of the workgroup's lines.  Data starts in `a`; returns the buffer
// holding the result (a or b).  Every thread of the workgroup must call it (it
// contains barriers); threads with active == false do no work.
template <bool INV>
__device__ float2 *lds_fft(float2 *a, float2 *b, const float2 *tw, const FftPlan &p, int LS,
                           int line, int jt, int tpl, bool active) {
    int Ns = 1;
    for (int s = 0; s < p.nst; s++) {
        const int R = p.radix[s];
        const float2 *t = tw + p.twoff[s];
        if (active) {
            switch (R) {
                case 2: fft_stage<2, INV>(a, b, t, p.n, Ns, LS, line, jt, tpl); break;
                case 3: fft_stage<3, INV>(a, b, t, p.n, Ns, LS, line, jt, tpl); break;
                case 4: fft_stage<4, INV>(a, b, t, p.n, Ns, LS, line, jt, tpl); break;
                case 5: fft_stage<5, INV>(a, b, t, p.n, Ns, LS, line, jt, tpl); break;
                case 7: fft_stage<7, INV>(a, b, t, p.n, Ns, LS, line, jt, tpl); break;
                case 8: fft_stage<8, INV>(a, b, t, p.n, Ns, LS, line, jt, tpl); break;
                case 11: fft_stage<11, INV>(a, b, t, p.n, Ns, LS, line, jt, tpl); break;
                case 13: fft_stage<13, INV>(a, b, t, p.n, Ns, LS, line, jt, tpl); break;
                default: break;
            }
        }
        __syncthreads();
        float2 *tmp = a;
        a = b;
        b = tmp;
        Ns *= R;
    }
    return a;
}
